// tile_plans.cpp -- host-side plans of the native kernels (no device code,
// no runtime calls): nnz tiles with split rows (k_nnz_tiled), SpMV tiles
// (k_spmv_tiled), column-cached tiles (k_nnz_cached) and strided row blocks
// (k_row_blocked).  api_native.cpp uploads them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

#include "matrix.hpp"
#include "native.hpp"

namespace dspmm {

// Cuts rows [0, nr) into tiles of <= S nonzeros and <= S rows that start and
// end on row boundaries; a row longer than S becomes ceil(len/S) chunks of
// S nonzeros, each with its own carry slot, and one split record.
void build_tiles(const std::vector<int> &rp, int nr, int S, std::vector<int4> &tiles,
                 std::vector<int4> &splits, int &nslots) {
    tiles.clear();
    splits.clear();
    nslots = 0;
    int r = 0;
    while (r < nr) {
        const int len = rp[(size_t)r + 1] - rp[(size_t)r];
        if (len > S) {
            const int nch = (len + S - 1) / S;
            splits.push_back(make_int4(0, r, nslots, nch));
            for (int c = 0; c < nch; c++) {
                const int pb = rp[(size_t)r] + c * S;
                const int pe = std::min(pb + S, rp[(size_t)r + 1]);
                tiles.push_back(make_int4(r, -1 - (nslots + c), pb, pe));
            }
            nslots += nch;
            ++r;
        } else {
            const int rb = r, p0 = rp[(size_t)r];
            int re = r + 1;
            while (re < nr && re - rb < S && rp[(size_t)re + 1] - p0 <= S) ++re;
            tiles.push_back(make_int4(rb, re, p0, rp[(size_t)re]));
            r = re;
        }
    }
    // Split rows with the most chunks first: their combine is the longest,
    // so it starts first.  A row's own combine order does not depend on
    // its place in this list.
    std::stable_sort(splits.begin(), splits.end(),
                     [](const int4 &a, const int4 &b) { return a.w > b.w; });
}

void build_spmv_tiles(const std::vector<int> &rp, int nr, int T, int row_max,
                      std::vector<int4> &tiles, std::vector<int4> &splits, int &nslots) {
    tiles.clear();
    splits.clear();
    nslots = 0;
    // Rows longer than row_max but at most kSpmvLongTiles * T get a tile of
    // their own; they are listed first, so the longest work starts first.
    std::vector<int4> longs;
    int r = 0;
    while (r < nr) {
        const int len = rp[(size_t)r + 1] - rp[(size_t)r];
        if (len > row_max) {
            if (len <= T) {
                tiles.push_back(make_int4(r, r + 1, rp[(size_t)r], rp[(size_t)r + 1]));
            } else if (len <= kSpmvLongTiles * T) {
                longs.push_back(make_int4(r, r + 1, rp[(size_t)r], rp[(size_t)r + 1]));
            } else {
                const int nch = (len + T - 1) / T;
                splits.push_back(make_int4(0, r, nslots, nch));
                for (int c = 0; c < nch; c++) {
                    const int pb = rp[(size_t)r] + c * T;
                    tiles.push_back(make_int4(r, -1 - (nslots + c), pb, std::min(pb + T, rp[(size_t)r + 1])));
                }
                nslots += nch;
            }
            ++r;
        } else {
            const int rb = r, p0 = rp[(size_t)r];
            int re = r + 1;
            while (re < nr && re - rb < 256 && rp[(size_t)re + 1] - rp[(size_t)re] <= row_max &&
                   rp[(size_t)re + 1] - p0 <= T)
                ++re;
            tiles.push_back(make_int4(rb, re, p0, rp[(size_t)re]));
            r = re;
        }
    }
    tiles.insert(tiles.begin(), longs.begin(), longs.end());
}

void build_cached_tiles(const std::vector<int> &rp, const std::vector<int> &ci, int nr, int nc,
                        int umax, int smax, std::vector<int4> &tiles, std::vector<int> &ucols,
                        std::vector<uint16_t> &lidx, double &reuse) {
    tiles.clear();
    ucols.clear();
    lidx.assign(ci.size(), 0);
    std::vector<int> in_tile((size_t)nc, -1), pos((size_t)nc, 0), in_row((size_t)nc, -1);
    int cur = -1, cur_u = 0, cur_nnz = 0, cur_rows = 0;
    int64_t direct_nnz = 0;
    auto fresh_cols = [&](int r, int tile) {           // distinct columns of r not in `tile`
        int f = 0;
        for (int p = rp[(size_t)r]; p < rp[(size_t)r + 1]; p++) {
            const int c = ci[(size_t)p];
            if (in_tile[(size_t)c] != tile && in_row[(size_t)c] != r) {
                in_row[(size_t)c] = r;
                f++;
            }
        }
        for (int p = rp[(size_t)r]; p < rp[(size_t)r + 1]; p++) in_row[(size_t)ci[(size_t)p]] = -1;
        return f;
    };
    for (int r = 0; r < nr; r++) {
        const int len = rp[(size_t)r + 1] - rp[(size_t)r];
        int f = cur >= 0 ? fresh_cols(r, cur) : 0;
        if (cur < 0 || cur_u + f > umax || cur_nnz + len > smax || cur_rows == kCachedRowsMax) {
            cur = -1;
            f = fresh_cols(r, -2);
            if (f > umax || len > smax) {               // does not fit alone: direct tile
                tiles.push_back(make_int4(r, -1, rp[(size_t)r], (int)ucols.size()));
                direct_nnz += len;
                continue;
            }
            cur = (int)tiles.size();
            tiles.push_back(make_int4(r, r, rp[(size_t)r], (int)ucols.size()));
            cur_u = cur_nnz = cur_rows = 0;
        }
        const int u0 = tiles[(size_t)cur].w;
        for (int p = rp[(size_t)r]; p < rp[(size_t)r + 1]; p++) {
            const int c = ci[(size_t)p];
            if (in_tile[(size_t)c] != cur) {
                in_tile[(size_t)c] = cur;
                pos[(size_t)c] = (int)ucols.size() - u0;
                ucols.push_back(c);
            }
            lidx[(size_t)p] = (uint16_t)pos[(size_t)c];
        }
        cur_u += f;
        cur_nnz += len;
        cur_rows++;
        tiles[(size_t)cur].y = r + 1;
    }
    tiles.push_back(make_int4(nr, nr, ci.empty() ? 0 : (int)ci.size(), (int)ucols.size()));
    const double reads = (double)ucols.size() + (double)direct_nnz;
    reuse = reads > 0 ? (double)ci.size() / reads : 1.0;
}

bool build_row_blocks(const std::vector<int> &rp, const std::vector<int> &ci,
                      const std::vector<float> &va, int nr, int R, int stride,
                      std::vector<int> &boff, std::vector<int2> &list, std::vector<float> &vals) {
    if (R < 1 || R > 8 || stride < 1) return false;    // a row's cursor per block row, 8 at most
    const int64_t period = (int64_t)R * stride;
    const int64_t nb64 = (nr + period - 1) / period * stride;
    if (nb64 >= INT_MAX) return false;
    const int nb = (int)nb64;
    boff.assign((size_t)nb + 1, 0);
    list.clear();
    vals.clear();
    for (int r = 0; r < nr; r++)                     // sorted rows without repeats
        for (int p = rp[(size_t)r] + 1; p < rp[(size_t)r + 1]; p++)
            if (ci[(size_t)p] <= ci[(size_t)p - 1]) return false;
    int head[8], end[8];
    for (int b = 0; b < nb; b++) {
        const int64_t r0 = (int64_t)(b / stride) * period + b % stride;
        const int nrows = r0 >= nr ? 0 : (int)std::min<int64_t>(R, (nr - 1 - r0) / stride + 1);
        for (int i = 0; i < nrows; i++) {
            head[i] = rp[(size_t)(r0 + (int64_t)i * stride)];
            end[i] = rp[(size_t)(r0 + (int64_t)i * stride) + 1];
        }
        for (;;) {
            int c = INT_MAX;
            for (int i = 0; i < nrows; i++)
                if (head[i] < end[i]) c = std::min(c, ci[(size_t)head[i]]);
            if (c == INT_MAX) break;
            int mask = 0;
            const size_t v0 = vals.size();
            vals.resize(v0 + (size_t)R, 0.f);        // R values per listed column, 0 where absent
            for (int i = 0; i < nrows; i++)
                if (head[i] < end[i] && ci[(size_t)head[i]] == c) {
                    mask |= 1 << i;
                    vals[v0 + (size_t)i] = va[(size_t)head[i]];
                    head[i]++;
                }
            list.push_back(make_int2(c, mask));
        }
        boff[(size_t)b + 1] = (int)list.size();
    }
    return true;
}

}  // namespace dspmm
