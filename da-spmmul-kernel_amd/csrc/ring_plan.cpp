// ring_plan.cpp -- host plan of the B-row ring kernel (k_ring): the order in
// which a workgroup visits rows, and a software-managed cache of B rows in
// LDS whose fills and evictions are decided here, once per matrix.
//
// Rows are grouped into tiles and the tiles into streams; one workgroup
// walks one stream.  With row distances s_mid < s_big (a grid stencil's
// line and plane length, found by sampling) a tile is a w x h block of
// rows -- w consecutive rows, h runs s_mid apart -- and a stream visits
// the same block in successive s_big-planes, so a tile's rows need mostly
// the B rows its predecessor already brought in (a 27-point stencil's
// tile needs three planes of its halo, two of them resident).  Without
// such distances a tile is w * h consecutive rows and a stream a run of
// consecutive tiles.
//
// Per stream the ring (U slots, one B row each) is simulated with Belady's
// rule: a B row a tile needs and the ring lacks is loaded into a free slot
// or into the slot of the resident row whose next use in the stream is
// farthest away (never used again first), among the rows that neither
// this tile nor the one before it reads: the kernel writes a tile's fills
// while its waves may still be computing the previous tile.  The ring
// holds, by default, the most distinct B rows two consecutive tiles of a
// stream need (rounded up to 8).  The plan lists, per tile, the
// (column, slot) fills to perform before the tile runs, and gives every
// nonzero the slot its B row occupies when its tile runs.  Each row keeps
// its nonzeros in storage order, so the kernel sums every row exactly like
// the host loop.  Every row of a tile is padded to the same length -- the
// tile's longest row rounded up to a multiple of 4 -- with (zero slot, 0)
// entries: the slot past the ring's last holds zeros, so a padding entry
// adds 0 * 0 = +0 to a running sum, which leaves it unchanged (an FP32 sum
// that starts at +0 is never -0).  The kernel's loops then have no bounds
// per row, and a lane group sums its rows of a tile side by side.
#include <algorithm>
#include <climits>
#include <numeric>
#include <vector>

#include "matrix.hpp"
#include "ring.hpp"

namespace dspmm {

namespace {

// Tiles of the plan in stream order: rows of each tile, and the streams.
struct Order {
    std::vector<int> rows;        // row ids, tile after tile
    std::vector<int> tile_off;    // ntiles + 1 offsets into rows
    std::vector<int> stream_off;  // nstreams + 1 offsets into tiles
};

void order_blocks(int nr, int w, int h, int s_mid, int s_big, int zchunk, Order &o) {
    // coordinates r = a + s_mid * b + s_big * c, a < s_mid, b < s_big / s_mid
    const int na = s_mid, nb = s_big / s_mid;
    const int nc = (int)((nr + (int64_t)s_big - 1) / s_big);
    const int bi = (na + w - 1) / w, bj = (nb + h - 1) / h;
    const int nz = (nc + zchunk - 1) / zchunk;
    o.rows.clear();
    o.tile_off.assign(1, 0);
    o.stream_off.assign(1, 0);
    for (int zc = 0; zc < nz; zc++)
        for (int j = 0; j < bj; j++)
            for (int i = 0; i < bi; i++) {
                for (int c = zc * zchunk; c < std::min(nc, (zc + 1) * zchunk); c++) {
                    const size_t before = o.rows.size();
                    for (int b = j * h; b < std::min(nb, (j + 1) * h); b++)
                        for (int a = i * w; a < std::min(na, (i + 1) * w); a++) {
                            const int64_t r = a + (int64_t)s_mid * b + (int64_t)s_big * c;
                            if (r < nr) o.rows.push_back((int)r);
                        }
                    if (o.rows.size() > before) o.tile_off.push_back((int)o.rows.size());
                }
                if ((int)o.tile_off.size() - 1 > o.stream_off.back())
                    o.stream_off.push_back((int)o.tile_off.size() - 1);
            }
}

void order_consecutive(int nr, int T, int tiles_per_stream, Order &o) {
    o.rows.resize((size_t)nr);
    std::iota(o.rows.begin(), o.rows.end(), 0);
    o.tile_off.assign(1, 0);
    for (int r = 0; r < nr; r += T) o.tile_off.push_back(std::min(nr, r + T));
    const int ntiles = (int)o.tile_off.size() - 1;
    o.stream_off.assign(1, 0);
    for (int t = 0; t < ntiles; t += tiles_per_stream)
        o.stream_off.push_back(std::min(ntiles, t + tiles_per_stream));
}

// Whether a shape walks (w x h)-blocks of a grid's planes (else consecutive
// tiles).
bool blocked(const RingShape &sh) {
    return sh.s_big > 1 && sh.s_mid >= 1 && sh.s_big % sh.s_mid == 0 && sh.s_mid >= sh.w &&
           (sh.h == 1 || sh.s_mid > 1);
}

}  // namespace

int ring_auto_chunk(int nr, const RingShape &sh, int cus) {
    if (nr <= 0 || sh.w < 1 || sh.h < 1) return 1;
    // walks side by side (one per block of a plane) and tiles along a walk
    int64_t cols = 1, planes;
    if (blocked(sh)) {
        const int64_t na = sh.s_mid, nb = sh.s_big / sh.s_mid;
        cols = ((na + sh.w - 1) / sh.w) * ((nb + sh.h - 1) / sh.h);
        planes = ((int64_t)nr + sh.s_big - 1) / sh.s_big;
    } else {
        planes = ((int64_t)nr + (int64_t)sh.w * sh.h - 1) / ((int64_t)sh.w * sh.h);
    }
    const int64_t n = std::max(1, cus);
    int best = 1;
    int64_t best_cost = INT64_MAX;
    for (int64_t z = std::min<int64_t>(planes, kRingMaxChunk); z >= 1; z--) {
        const int64_t walks = cols * ((planes + z - 1) / z);
        const int64_t cost = ((walks + n - 1) / n) * (kRingFirstTileCost + z);
        if (cost < best_cost) {
            best_cost = cost;
            best = (int)z;
        }
    }
    return best;
}

bool build_ring_plan(const std::vector<int> &rp, const std::vector<int> &ci,
                     const std::vector<float> &va, int nr, int nc, const RingShape &sh,
                     RingPlanHost &p) {
    p = RingPlanHost{};
    if (nr <= 0 || sh.w < 1 || sh.h < 1 || sh.slots < 0 || sh.slots > 65535) return false;
    Order o;
    if (blocked(sh))
        order_blocks(nr, sh.w, sh.h, sh.s_mid, sh.s_big, std::max(1, sh.zchunk), o);
    else
        order_consecutive(nr, sh.w * sh.h, std::max(1, sh.zchunk), o);
    const int ntiles = (int)o.tile_off.size() - 1;
    const int nstreams = (int)o.stream_off.size() - 1;
    // every row exactly once
    {
        std::vector<char> seen((size_t)nr, 0);
        for (int r : o.rows) {
            if (r < 0 || r >= nr || seen[(size_t)r]) return false;
            seen[(size_t)r] = 1;
        }
        if ((int)o.rows.size() != nr) return false;
    }
    p.rows_max = 0;
    for (int t = 0; t < ntiles; t++)
        p.rows_max = std::max(p.rows_max, o.tile_off[(size_t)t + 1] - o.tile_off[(size_t)t]);

    // per tile: its distinct columns (sorted)
    std::vector<int> need_off((size_t)ntiles + 1, 0), need;
    int need_max = 0;
    {
        std::vector<int> buf;
        for (int t = 0; t < ntiles; t++) {
            buf.clear();
            for (int q = o.tile_off[(size_t)t]; q < o.tile_off[(size_t)t + 1]; q++) {
                const int r = o.rows[(size_t)q];
                buf.insert(buf.end(), ci.begin() + rp[(size_t)r], ci.begin() + rp[(size_t)r + 1]);
            }
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            need_max = std::max(need_max, (int)buf.size());
            need.insert(need.end(), buf.begin(), buf.end());
            need_off[(size_t)t + 1] = (int)need.size();
        }
    }
    // the most distinct columns of two consecutive tiles of a stream
    int pair_max = need_max;
    for (int s = 0; s < nstreams; s++)
        for (int t = o.stream_off[(size_t)s] + 1; t < o.stream_off[(size_t)s + 1]; t++) {
            const auto a0 = need.begin() + need_off[(size_t)t - 1], a1 = need.begin() + need_off[(size_t)t];
            const auto b1 = need.begin() + need_off[(size_t)t + 1];
            // |A u B| = |A| + |B| - |A n B| over the sorted sets
            int common = 0;
            for (auto x = a0, y = a1; x != a1 && y != b1;) {
                if (*x < *y) ++x;
                else if (*y < *x) ++y;
                else { ++common; ++x; ++y; }
            }
            pair_max = std::max(pair_max, (int)(a1 - a0) + (int)(b1 - a1) - common);
        }
    const int slots = sh.slots > 0 ? sh.slots : std::max(8, (pair_max + 7) / 8 * 8);
    if (pair_max > slots || slots > 65534) return false;   // two consecutive tiles must fit
    p.slots = slots;

    // Belady simulation, stream by stream
    std::vector<int> slot_of_col((size_t)nc, -1);        // reset per stream for touched columns
    std::vector<int> occ_head((size_t)nc, -1);           // per column: next occurrence index
    std::vector<int> occ_next(need.size(), -1);          // chain of a column's occurrences
    std::vector<int> occ_tile(need.size(), 0);
    std::vector<int> col_of_slot((size_t)slots, -1);
    std::vector<int> needed_at((size_t)nc, -1);          // last tile that needs the column
    std::vector<std::pair<int, int>> cand;               // (next use, slot)
    std::vector<int> tile_slot(need.size(), -1);          // slot of need[i] when its tile runs
    p.tiles.resize((size_t)ntiles);
    p.fill_off.assign((size_t)ntiles + 1, 0);
    for (int s = 0; s < nstreams; s++) {
        const int t0 = o.stream_off[(size_t)s], t1 = o.stream_off[(size_t)s + 1];
        // occurrence chains (built backwards so that heads are the earliest)
        for (int t = t1 - 1; t >= t0; t--)
            for (int i = need_off[(size_t)t]; i < need_off[(size_t)t + 1]; i++) {
                const int c = need[(size_t)i];
                occ_next[(size_t)i] = occ_head[(size_t)c];
                occ_tile[(size_t)i] = t;
                occ_head[(size_t)c] = i;
            }
        std::fill(col_of_slot.begin(), col_of_slot.end(), -1);
        int free_next = 0;
        for (int t = t0; t < t1; t++) {
            // advance each needed column's chain past this tile
            std::vector<int> missing;
            for (int i = need_off[(size_t)t]; i < need_off[(size_t)t + 1]; i++) {
                const int c = need[(size_t)i];
                occ_head[(size_t)c] = occ_next[(size_t)i];     // next use after t
                needed_at[(size_t)c] = t;
                if (slot_of_col[(size_t)c] < 0) missing.push_back(i);
            }
            // victims: free slots first, then the farthest next use among
            // residents neither this tile nor the previous one needs
            int nfree = slots - free_next;
            if ((int)missing.size() > nfree) {
                cand.clear();
                for (int u = 0; u < free_next; u++) {
                    const int c = col_of_slot[(size_t)u];
                    if (c < 0) { cand.emplace_back(INT_MAX, u); continue; }
                    if (needed_at[(size_t)c] >= std::max(t - 1, t0)) continue;  // read by t or t - 1
                    const int h = occ_head[(size_t)c];
                    cand.emplace_back(h < 0 ? INT_MAX : occ_tile[(size_t)h], u);
                }
                std::sort(cand.begin(), cand.end(), [](const auto &x, const auto &y) {
                    return x.first != y.first ? x.first > y.first : x.second < y.second;
                });
            }
            size_t ci_next = 0;
            for (int i : missing) {
                int u;
                if (free_next < slots) {
                    u = free_next++;
                } else {
                    if (ci_next >= cand.size()) return false;      // cannot happen: tile fits
                    u = cand[ci_next++].second;
                    const int old = col_of_slot[(size_t)u];
                    if (old >= 0) slot_of_col[(size_t)old] = -1;
                }
                const int c = need[(size_t)i];
                col_of_slot[(size_t)u] = c;
                slot_of_col[(size_t)c] = u;
                p.fills.push_back(make_int2(c, u));
            }
            p.fill_off[(size_t)t + 1] = (int)p.fills.size();
            for (int i = need_off[(size_t)t]; i < need_off[(size_t)t + 1]; i++)
                tile_slot[(size_t)i] = slot_of_col[(size_t)need[(size_t)i]];
            // columns whose chain ended are never used again in this stream:
            // their slots are candidates at distance INT_MAX (occ_head = -1)
        }
        // reset touched columns
        for (int t = t0; t < t1; t++)
            for (int i = need_off[(size_t)t]; i < need_off[(size_t)t + 1]; i++) {
                slot_of_col[(size_t)need[(size_t)i]] = -1;
                occ_head[(size_t)need[(size_t)i]] = -1;
                needed_at[(size_t)need[(size_t)i]] = -1;
            }
    }

    // nonzeros in tile-row order, each row padded to a multiple of 4
    // entries and each tile's range to a multiple of 8 (16-byte aligned
    // vector loads of values and slots)
    const uint16_t zero = (uint16_t)slots;
    p.zero_slot = slots;
    p.stream_off = o.stream_off;
    p.row_info.resize(o.rows.size());
    p.fill_max = 0;
    p.nnz_max = 0;
    size_t pos = 0;
    for (int t = 0; t < ntiles; t++) {
        const int q0 = o.tile_off[(size_t)t], q1 = o.tile_off[(size_t)t + 1];
        const int p0 = (int)pos;
        int rlen = 0;                                   // the tile's longest row, then padded
        for (int q = q0; q < q1; q++) {
            const int r = o.rows[(size_t)q];
            rlen = std::max(rlen, rp[(size_t)r + 1] - rp[(size_t)r]);
        }
        const int lmax = rlen;
        rlen = (rlen + 3) / 4 * 4;
        p.pad_entries += (int64_t)rlen * (q1 - q0);
        for (int q = q0; q < q1; q++) {
            const int r = o.rows[(size_t)q];
            for (int k = rp[(size_t)r]; k < rp[(size_t)r + 1]; k++) {
                const int c = ci[(size_t)k];
                // slot of c in this tile (binary search in the tile's sorted set)
                const auto b = need.begin() + need_off[(size_t)t], e = need.begin() + need_off[(size_t)t + 1];
                const auto it = std::lower_bound(b, e, c);
                p.pslot.push_back((uint16_t)tile_slot[(size_t)(it - need.begin())]);
                p.pval.push_back(va[(size_t)k]);
                pos++;
            }
            while ((int)pos - p0 < (q - q0 + 1) * rlen) {
                p.pslot.push_back(zero);
                p.pval.push_back(0.f);
                pos++;
            }
            p.row_info[(size_t)q] = make_int2(r, (int)pos - p0);     // row, end offset in tile
        }
        while (pos % 8) {
            p.pslot.push_back(zero);
            p.pval.push_back(0.f);
            pos++;
        }
        const int nf = p.fill_off[(size_t)t + 1] - p.fill_off[(size_t)t];
        // fills of a stream's first tile are done before the walk starts
        const bool first = std::binary_search(o.stream_off.begin(), o.stream_off.end(), t);
        if (!first) p.fill_max = std::max(p.fill_max, nf);
        p.nnz_max = std::max(p.nnz_max, (int)pos - p0);
        p.tiles[(size_t)t] = make_int4(p0, lmax, q0, q1);
    }
    p.total_fills = (int64_t)p.fills.size();
    for (int s = 0; s < nstreams; s++) {
        const int t0 = o.stream_off[(size_t)s], t1 = o.stream_off[(size_t)s + 1];
        p.stream_tiles_max = std::max(p.stream_tiles_max, t1 - t0);
        p.stream_fills_max = std::max(p.stream_fills_max,
                                      p.fill_off[(size_t)t1] - p.fill_off[(size_t)t0]);
    }
    // padding read (never used) by the kernel's branch-free prefetch: eight
    // entries past the last tile's nonzeros and one fill past the last
    for (int k = 0; k < 8; k++) {
        p.pslot.push_back(zero);
        p.pval.push_back(0.f);
    }
    p.fills.push_back(make_int2(0, 0));
    return true;
}

}  // namespace dspmm
