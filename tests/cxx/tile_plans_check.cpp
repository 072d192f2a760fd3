// Host-side check of the native kernels' plans (csrc/tile_plans.cpp), no
// GPU: for each planner and each matrix below, every row and nonzero is
// covered once and in order, and each tile keeps the limits its kernel
// relies on.  nnz tiles: <= S rows and nonzeros, split rows in S-chunks with
// consecutive carry slots.  SpMV tiles: rows <= row_max share tiles of <=
// 256 rows and T nonzeros, longer rows get tiles of their own (split past
// 16 T).  Column-cached tiles: the local index of every nonzero names its
// column, <= umax columns and <= smax nonzeros per tile, direct tiles only
// for rows that do not fit.  Row blocks: every row rebuilt from its blocks'
// column lists and values equals the CSR row bit for bit, and unsorted rows
// are refused.  Built once plainly and once with the host code under
// AddressSanitizer by tests/test_tile_plans.py.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <set>
#include <vector>

#include "matrix.hpp"
#include "native.hpp"

using namespace dspmm;

struct Csr {
    int nr = 0, nc = 0;
    std::vector<int> rp, ci;
    std::vector<float> va;
};

static uint64_t g_s = 1;
static uint64_t rnd() {
    g_s = g_s * 6364136223846793005ull + 1442695040888963407ull;
    return g_s >> 33;
}

// Rows of 0..maxlen sorted distinct columns; rows listed in `longs` get
// that many nonzeros instead (capped at nc).
static Csr make(int nr, int nc, int maxlen, uint64_t seed, std::vector<std::pair<int, int>> longs = {}) {
    g_s = seed;
    Csr a;
    a.nr = nr;
    a.nc = nc;
    a.rp.push_back(0);
    for (int r = 0; r < nr; r++) {
        int len = (int)(rnd() % (uint64_t)(maxlen + 1));
        for (auto &l : longs)
            if (l.first == r) len = l.second;
        len = std::min(len, nc);
        std::set<int> cols;
        while ((int)cols.size() < len) cols.insert((int)(rnd() % (uint64_t)nc));
        for (int c : cols) {
            a.ci.push_back(c);
            a.va.push_back((float)(rnd() % 2001) / 1000.f - 1.f);
        }
        a.rp.push_back((int)a.ci.size());
    }
    return a;
}

static int fail(const char *what, const char *name, int r) {
    std::printf("%s: %s (row %d)\n", name, what, r);
    return 1;
}

static int check_nnz_tiles(const char *name, const Csr &a, int S) {
    std::vector<int4> tiles, splits;
    int nslots = 0;
    build_tiles(a.rp, a.nr, S, tiles, splits, nslots);
    int r = 0, slot = 0;
    size_t t = 0;
    int nsplit_rows = 0;
    while (t < tiles.size()) {
        const int4 q = tiles[t];
        if (q.x != r) return fail("nnz tiles out of order", name, r);
        if (q.y >= 0) {
            if (q.y <= q.x || q.y - q.x > S || q.z != a.rp[(size_t)q.x] || q.w != a.rp[(size_t)q.y] ||
                q.w - q.z > S)
                return fail("nnz tile bounds", name, r);
            r = q.y;
            t++;
            continue;
        }
        const int len = a.rp[(size_t)r + 1] - a.rp[(size_t)r];
        const int nch = (len + S - 1) / S;
        if (len <= S) return fail("a short row split", name, r);
        for (int c = 0; c < nch; c++, t++) {
            const int pb = a.rp[(size_t)r] + c * S;
            if (t >= tiles.size() || tiles[t].x != r || tiles[t].y != -1 - (slot + c) || tiles[t].z != pb ||
                tiles[t].w != std::min(pb + S, a.rp[(size_t)r + 1]))
                return fail("split chunks", name, r);
        }
        bool found = false;
        for (const int4 &s : splits)
            if (s.y == r && s.z == slot && s.w == nch) found = true;
        if (!found) return fail("split record", name, r);
        slot += nch;
        nsplit_rows++;
        r++;
    }
    if (r != a.nr || slot != nslots || nsplit_rows != (int)splits.size()) return fail("nnz tile coverage", name, r);
    return 0;
}

static int check_spmv_tiles(const char *name, const Csr &a, int T, int row_max) {
    std::vector<int4> tiles, splits;
    int nslots = 0;
    build_spmv_tiles(a.rp, a.nr, T, row_max, tiles, splits, nslots);
    std::vector<int> seen((size_t)a.nr, 0);
    int slot = 0;
    size_t t = 0;
    bool longs_done = false;
    while (t < tiles.size()) {
        const int4 q = tiles[t];
        if (q.x < 0 || q.x >= a.nr) return fail("spmv tile row", name, q.x);
        const int len = a.rp[(size_t)q.x + 1] - a.rp[(size_t)q.x];
        if (q.y >= 0) {
            const bool own_long = q.y == q.x + 1 && len > T && len <= kSpmvLongTiles * T;
            if (own_long && longs_done) return fail("a long row's tile after the others", name, q.x);
            if (!own_long) longs_done = true;
            if (q.z != a.rp[(size_t)q.x] || q.w != a.rp[(size_t)q.y] || q.y - q.x > 256 ||
                (q.y - q.x > 1 && q.w - q.z > T))
                return fail("spmv tile bounds", name, q.x);
            for (int i = q.x; i < q.y; i++) {
                const int li = a.rp[(size_t)i + 1] - a.rp[(size_t)i];
                if (q.y - q.x > 1 && li > row_max) return fail("a long row in a shared tile", name, i);
                if (li > kSpmvLongTiles * T) return fail("an unsplit very long row", name, i);
                seen[(size_t)i]++;
            }
            t++;
            continue;
        }
        longs_done = true;
        const int nch = (len + T - 1) / T;
        if (len <= kSpmvLongTiles * T) return fail("a row split below 16 T", name, q.x);
        for (int c = 0; c < nch; c++, t++) {
            const int pb = a.rp[(size_t)q.x] + c * T;
            if (t >= tiles.size() || tiles[t].x != q.x || tiles[t].y != -1 - (slot + c) || tiles[t].z != pb ||
                tiles[t].w != std::min(pb + T, a.rp[(size_t)q.x + 1]))
                return fail("spmv split chunks", name, q.x);
        }
        slot += nch;
        seen[(size_t)q.x]++;
    }
    for (int i = 0; i < a.nr; i++)
        if (seen[(size_t)i] != 1) return fail("spmv row coverage", name, i);
    if (slot != nslots) return fail("spmv carry slots", name, -1);
    return 0;
}

static int check_cached_tiles(const char *name, const Csr &a, int umax, int smax) {
    std::vector<int4> tiles;
    std::vector<int> ucols;
    std::vector<uint16_t> lidx;
    double reuse = 0;
    build_cached_tiles(a.rp, a.ci, a.nr, a.nc, umax, smax, tiles, ucols, lidx, reuse);
    if (tiles.empty()) return fail("no sentinel", name, -1);
    const int4 sent = tiles.back();
    if (sent.x != a.nr || sent.w != (int)ucols.size()) return fail("sentinel", name, -1);
    int r = 0;
    for (size_t t = 0; t + 1 < tiles.size(); t++) {
        const int4 q = tiles[t];
        const int u0 = q.w, u1 = tiles[t + 1].w;
        if (q.x != r || q.z != a.rp[(size_t)r]) return fail("cached tiles out of order", name, r);
        if (q.y < 0) {
            const int len = a.rp[(size_t)r + 1] - a.rp[(size_t)r];
            if (len <= smax && len <= umax) return fail("a direct tile for a row that fits", name, r);
            if (u1 != u0) return fail("a direct tile lists columns", name, r);
            r++;
            continue;
        }
        if (q.y <= q.x || q.y - q.x > kCachedRowsMax || u1 - u0 > umax ||
            a.rp[(size_t)q.y] - a.rp[(size_t)q.x] > smax)
            return fail("cached tile limits", name, r);
        std::set<int> distinct(ucols.begin() + u0, ucols.begin() + u1);
        if ((int)distinct.size() != u1 - u0) return fail("a column listed twice", name, r);
        for (int p = a.rp[(size_t)q.x]; p < a.rp[(size_t)q.y]; p++)
            if ((int)lidx[(size_t)p] >= u1 - u0 || ucols[(size_t)u0 + lidx[(size_t)p]] != a.ci[(size_t)p])
                return fail("local index", name, r);
        r = q.y;
    }
    if (r != a.nr) return fail("cached row coverage", name, r);
    return 0;
}

static int check_row_blocks(const char *name, const Csr &a, int R, int stride) {
    std::vector<int> boff;
    std::vector<int2> list;
    std::vector<float> vals;
    if (!build_row_blocks(a.rp, a.ci, a.va, a.nr, R, stride, boff, list, vals))
        return fail("row blocks refused", name, -1);
    if (vals.size() != list.size() * (size_t)R || boff.back() != (int)list.size())
        return fail("row block sizes", name, -1);
    std::vector<std::vector<std::pair<int, float>>> rows((size_t)a.nr);
    const int nb = (int)boff.size() - 1;
    for (int b = 0; b < nb; b++) {
        const int64_t r0 = (int64_t)(b / stride) * R * stride + b % stride;
        for (int e = boff[(size_t)b]; e < boff[(size_t)b + 1]; e++) {
            if (e > boff[(size_t)b] && list[(size_t)e].x <= list[(size_t)e - 1].x)
                return fail("block columns not ascending", name, (int)r0);
            for (int i = 0; i < R; i++) {
                const float v = vals[(size_t)e * R + i];
                if (list[(size_t)e].y >> i & 1) {
                    const int64_t r = r0 + (int64_t)i * stride;
                    if (r >= a.nr) return fail("a mask bit past the last row", name, (int)r);
                    rows[(size_t)r].push_back({list[(size_t)e].x, v});
                } else if (v != 0.f) {
                    return fail("a value where the mask is clear", name, (int)r0);
                }
            }
        }
    }
    for (int r = 0; r < a.nr; r++) {
        const int p0 = a.rp[(size_t)r], len = a.rp[(size_t)r + 1] - p0;
        if ((int)rows[(size_t)r].size() != len) return fail("row length through the blocks", name, r);
        for (int k = 0; k < len; k++)
            if (rows[(size_t)r][(size_t)k].first != a.ci[(size_t)p0 + k] ||
                std::memcmp(&rows[(size_t)r][(size_t)k].second, &a.va[(size_t)p0 + k], 4) != 0)
                return fail("row through the blocks differs", name, r);
    }
    return 0;
}

int main() {
    int rc = 0, ok = 0;
    struct M { const char *name; Csr a; };
    std::vector<M> mats;
    mats.push_back({"random 3000 rows", make(3000, 5000, 40, 11)});
    mats.push_back({"long rows", make(2000, 100000, 12, 13, {{3, 5000}, {700, 70000}, {1999, 300}})});
    mats.push_back({"empty rows only", make(50, 10, 0, 15)});
    mats.push_back({"one row", make(1, 30, 25, 17)});
    mats.push_back({"no rows", make(0, 10, 5, 19)});
    mats.push_back({"dense small", make(300, 64, 64, 21)});
    for (const auto &m : mats) {
        int e = 0;
        for (int S : {16, 64, 256}) e |= check_nnz_tiles(m.name, m.a, S);
        for (int T : {512, 4096}) e |= check_spmv_tiles(m.name, m.a, T, 16);
        for (int u : {32, 256}) e |= check_cached_tiles(m.name, m.a, u, std::min(4096, 4 * u));
        for (int R : {2, 4, 8})
            for (int s : {1, 3, 100}) e |= check_row_blocks(m.name, m.a, R, s);
        if (!e) {
            std::printf("%s: ok\n", m.name);
            ok++;
        }
        rc |= e;
    }
    {   // an unsorted row, and a block height past the planner's limit, are refused
        Csr u = make(10, 20, 5, 23);
        std::swap(u.ci[(size_t)u.rp[2]], u.ci[(size_t)u.rp[3] - 1]);
        std::vector<int> boff;
        std::vector<int2> list;
        std::vector<float> vals;
        const bool unsorted_ok = u.rp[3] - u.rp[2] >= 2 &&
                                 build_row_blocks(u.rp, u.ci, u.va, u.nr, 4, 1, boff, list, vals);
        const bool tall_ok = build_row_blocks(mats[0].a.rp, mats[0].a.ci, mats[0].a.va, mats[0].a.nr, 16, 1,
                                              boff, list, vals);
        if (unsorted_ok || tall_ok) {
            std::printf("refusals: unsorted accepted %d, 16-row blocks accepted %d\n", unsorted_ok, tall_ok);
            rc |= 1;
        } else {
            std::printf("refusals: ok\n");
            ok++;
        }
    }
    std::printf("%d of %d checks ok\n", ok, (int)mats.size() + 1);
    return rc;
}
