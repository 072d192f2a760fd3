// Host-side check of the B-row ring plan (csrc/ring_plan.cpp), no GPU:
// replays every stream's fills into a simulated ring and checks that each
// nonzero's slot holds its column when its tile runs, that no fill of a
// tile overwrites a slot the previous tile of its stream reads (the kernel
// writes a tile's fills while the previous tile may still be computing),
// that the plan's values and row ends follow the CSR arrays, that every row
// appears once, and that the limits the kernel relies on hold (slots <
// capacity, tiles' nonzero ranges 8-aligned and inside the arrays).  Then it evaluates C through the plan in the
// kernel's order and compares it bit for bit with the row-by-row host loop.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ring.hpp"

using dspmm::RingPlanHost;
using dspmm::RingShape;

struct Csr {
    int nr = 0, nc = 0;
    std::vector<int> rp, ci;
    std::vector<float> va;
};

static float hval(uint64_t i) {
    uint64_t z = i * 0x9E3779B97F4A7C15ull + 12345;
    z = (z ^ (z >> 31)) * 0xBF58476D1CE4E5B9ull;
    return (float)((z >> 40) & 0xffff) / 65536.0f - 0.5f;
}

// 27-point stencil on an nx x ny x nz grid, x fastest.
static Csr stencil3(int nx, int ny, int nz) {
    Csr a;
    a.nr = a.nc = nx * ny * nz;
    a.rp.push_back(0);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int xx = x + dx, yy = y + dy, zz = z + dz;
                            if (xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
                            a.ci.push_back(xx + nx * (yy + ny * zz));
                            a.va.push_back(hval(a.ci.size()));
                        }
                a.rp.push_back((int)a.ci.size());
            }
    return a;
}

static Csr stencil(int n) { return stencil3(n, n, n); }

// 7-point stencil on an nx x ny x nz grid (nz = 1: the 5-point stencil).
static Csr stencil7(int nx, int ny, int nz) {
    Csr a;
    a.nr = a.nc = nx * ny * nz;
    a.rp.push_back(0);
    const int d[7][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                for (const auto &o : d) {
                    const int xx = x + o[0], yy = y + o[1], zz = z + o[2];
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
                    a.ci.push_back(xx + nx * (yy + ny * zz));
                    a.va.push_back(hval(a.ci.size() + 5));
                }
                a.rp.push_back((int)a.ci.size());
            }
    return a;
}

// Random rows of 0..maxlen sorted distinct columns (some empty).
static Csr random_rows(int nr, int nc, int maxlen, uint64_t seed) {
    Csr a;
    a.nr = nr;
    a.nc = nc;
    a.rp.push_back(0);
    uint64_t s = seed;
    auto next = [&]() { s = s * 6364136223846793005ull + 1442695040888963407ull; return s >> 33; };
    std::vector<char> used((size_t)nc, 0);
    for (int r = 0; r < nr; r++) {
        const int len = (int)(next() % (uint64_t)(maxlen + 1));
        std::vector<int> cols;
        for (int k = 0; k < len; k++) {
            const int c = (int)(next() % (uint64_t)nc);
            if (!used[(size_t)c]) { used[(size_t)c] = 1; cols.push_back(c); }
        }
        std::sort(cols.begin(), cols.end());
        for (int c : cols) {
            used[(size_t)c] = 0;
            a.ci.push_back(c);
            a.va.push_back(hval(a.ci.size() + seed));
        }
        a.rp.push_back((int)a.ci.size());
    }
    return a;
}

static int check(const char *name, const Csr &a, const RingShape &sh, int nv) {
    RingPlanHost p;
    if (!dspmm::build_ring_plan(a.rp, a.ci, a.va, a.nr, a.nc, sh, p)) {
        std::printf("%s: plan refused\n", name);
        return 1;
    }
    const int U = p.slots;
    const int ntiles = (int)p.tiles.size();
    std::vector<char> seen((size_t)a.nr, 0);
    // B and the host loop's C
    std::vector<float> B((size_t)a.nc * nv), Chost((size_t)a.nr * nv, 0.f), Cplan((size_t)a.nr * nv, -1.f);
    for (size_t i = 0; i < B.size(); i++) B[i] = hval(i * 7 + 3);
    for (int r = 0; r < a.nr; r++)
        for (int k = 0; k < nv; k++) {
            float acc = 0.f;
            for (int q = a.rp[(size_t)r]; q < a.rp[(size_t)r + 1]; q++)
                acc = acc + a.va[(size_t)q] * B[(size_t)a.ci[(size_t)q] * nv + k];
            Chost[(size_t)r * nv + k] = acc;
        }
    int64_t fills = 0;
    for (size_t st = 0; st + 1 < p.stream_off.size(); st++) {
        std::vector<int> ring((size_t)U, -1);
        std::vector<int> last_read((size_t)U, INT_MIN); // last tile of the stream reading a slot
        for (int t = p.stream_off[st]; t < p.stream_off[st + 1]; t++) {
            for (int f = p.fill_off[(size_t)t]; f < p.fill_off[(size_t)t + 1]; f++) {
                const int2 fl = p.fills[(size_t)f];
                if (fl.y < 0 || fl.y >= U || fl.x < 0 || fl.x >= a.nc) {
                    std::printf("%s: fill out of range\n", name);
                    return 2;
                }
                if (last_read[(size_t)fl.y] >= t - 1) {
                    std::printf("%s: tile %d fills slot %d, which tile %d reads\n", name, t, fl.y,
                                last_read[(size_t)fl.y]);
                    return 11;
                }
                ring[(size_t)fl.y] = fl.x;
                fills++;
            }
            const int4 td = p.tiles[(size_t)t];
            // td.y: the longest row; the tile's entries: rows x (td.y rounded
            // to 4), rounded to 8
            const int ntr = td.w - td.z, plen = (td.y + 3) / 4 * 4;
            const int span = (ntr * plen + 7) / 8 * 8;
            // the kernel loads whole 8-entry chunks: the last one must lie
            // inside the arrays too
            if (td.x % 8 || span > p.nnz_max || ntr > p.rows_max ||
                (size_t)td.x + (size_t)span > p.pval.size() ||
                (size_t)td.x + (size_t)span > p.pslot.size() ||
                (t + 1 < (int)p.tiles.size() && p.tiles[(size_t)t + 1].x < td.x + span)) {
                std::printf("%s: tile %d range\n", name, t);
                return 3;
            }
            int longest = 0;
            for (int q = td.z; q < td.w; q++) {
                const int r = p.row_info[(size_t)q].x;
                longest = std::max(longest, a.rp[(size_t)r + 1] - a.rp[(size_t)r]);
            }
            if (longest != td.y) {
                std::printf("%s: tile %d longest row %d, descriptor says %d\n", name, t, longest, td.y);
                return 12;
            }
            int prev = 0;
            for (int q = td.z; q < td.w; q++) {
                const int2 ri = p.row_info[(size_t)q];
                const int r = ri.x;
                if (r < 0 || r >= a.nr || seen[(size_t)r]) { std::printf("%s: row %d\n", name, r); return 4; }
                seen[(size_t)r] = 1;
                const int len = a.rp[(size_t)r + 1] - a.rp[(size_t)r];
                const int padded = ri.y - prev;
                // every row of the tile padded to the same length: the longest
                // row rounded up to 4 (the kernel derives it from the descriptor)
                const int first = p.row_info[(size_t)td.z].y;
                if (padded % 4 || padded < len || padded != first || padded != plen || ri.y > span) {
                    std::printf("%s: row %d end\n", name, r);
                    return 5;
                }
                for (int k = 0; k < nv; k++) {
                    // the kernel's order: the row's nonzeros, then its
                    // padding (zero slot, value 0) -- 0 * 0 added to the sum
                    float acc = 0.f;
                    for (int e = 0; e < padded; e++) {
                        const int pe = td.x + prev + e;
                        const int slot = p.pslot[(size_t)pe];
                        if (e >= len) {
                            if (slot != p.zero_slot || p.pval[(size_t)pe] != 0.f) {
                                std::printf("%s: row %d padding entry %d\n", name, r, e);
                                return 10;
                            }
                            acc = acc + p.pval[(size_t)pe] * 0.f;
                            continue;
                        }
                        const int col = a.ci[(size_t)a.rp[(size_t)r] + e];
                        if (slot >= U || ring[(size_t)slot] != col || p.pval[(size_t)pe] != a.va[(size_t)a.rp[(size_t)r] + e]) {
                            std::printf("%s: tile %d row %d nonzero %d: slot %d holds %d, wants %d\n", name, t,
                                        r, e, slot, slot < U ? ring[(size_t)slot] : -9, col);
                            return 6;
                        }
                        acc = acc + p.pval[(size_t)pe] * B[(size_t)ring[(size_t)slot] * nv + k];
                        last_read[(size_t)slot] = t;
                    }
                    Cplan[(size_t)r * nv + k] = acc;
                }
                prev = ri.y;
            }
        }
    }
    for (int r = 0; r < a.nr; r++)
        if (!seen[(size_t)r]) { std::printf("%s: row %d missing\n", name, r); return 7; }
    if (std::memcmp(Chost.data(), Cplan.data(), sizeof(float) * Chost.size()) != 0) {
        std::printf("%s: C through the plan differs from the host loop\n", name);
        return 8;
    }
    if (fills != p.total_fills || p.zero_slot != U) return 9;
    std::printf("%s: ok, %d streams, %d tiles, %d slots, %.3f fills per row, nnz_max %d, rows_max %d\n",
                name, (int)p.stream_off.size() - 1, ntiles, U, (double)fills / std::max(1, a.nr),
                p.nnz_max, p.rows_max);
    return 0;
}

int main() {
    int rc = 0;
    {
        const int n = 21;                      // w does not divide the line length
        Csr a = stencil(n);
        for (int w : {8, 4, 16})
            for (int h : {4, 8})
                for (int z : {3, 10}) {
                    RingShape sh;
                    sh.w = w;
                    sh.h = h;
                    sh.s_mid = n;
                    sh.s_big = n * n;
                    sh.zchunk = z;
                    char name[64];
                    std::snprintf(name, sizeof name, "stencil%d w%d h%d z%d", n, w, h, z);
                    rc |= check(name, a, sh, 8);
                }
        // a fixed, larger ring: Belady keeps more than the halo
        RingShape sh;
        sh.s_mid = n;
        sh.s_big = n * n;
        sh.slots = 400;
        rc |= check("stencil21 w8 h4 z25 U400", a, sh, 4);
        // one distance only: tiles of consecutive rows walked along it
        RingShape s1;
        s1.w = 16;
        s1.h = 1;
        s1.s_mid = n * n;
        s1.s_big = n * n;
        s1.zchunk = 7;
        rc |= check("stencil21 one distance", a, s1, 4);
    }
    {
        Csr a = random_rows(3000, 5000, 40, 11);     // no structure: consecutive tiles
        RingShape sh;
        sh.w = 8;
        sh.h = 4;
        sh.zchunk = 5;
        rc |= check("random consecutive", a, sh, 4);
        sh.slots = 2048;                             // roomy ring
        rc |= check("random consecutive U2048", a, sh, 4);
    }
    {
        Csr a = stencil(12);
        RingShape sh;                                // more slots than a tile needs
        sh.s_mid = 12;
        sh.s_big = 144;
        sh.w = 4;
        sh.h = 4;
        sh.zchunk = 12;
        rc |= check("stencil12 w4 h4", a, sh, 16);
        sh.slots = 40;                               // fewer than a tile needs: refused
        RingPlanHost p;
        if (dspmm::build_ring_plan(a.rp, a.ci, a.va, a.nr, a.nc, sh, p)) {
            std::printf("a ring smaller than a tile's columns must be refused\n");
            rc |= 1;
        }
    }
    {
        // grids smaller than one tile, a grid whose line and plane lengths
        // no tile dimension divides (partial tiles on every edge), and the
        // shapes the library picks for it
        struct { int nx, ny, nz, w, h, z; } g[] = {
            {2, 2, 2, 4, 8, 0}, {3, 3, 3, 4, 8, 0}, {3, 3, 3, 8, 8, 5}, {1, 1, 5, 4, 4, 3},
            {7, 9, 11, 4, 8, 0}, {7, 9, 11, 8, 4, 4}, {13, 5, 3, 8, 8, 17}, {37, 2, 2, 16, 4, 2}};
        for (const auto &q : g) {
            Csr a = stencil3(q.nx, q.ny, q.nz);
            RingShape sh;
            sh.w = q.w;
            sh.h = q.h;
            sh.s_mid = q.nx;
            sh.s_big = q.nx * q.ny;
            sh.zchunk = q.z > 0 ? q.z : dspmm::ring_auto_chunk(a.nr, sh, 256);
            char name[96];
            std::snprintf(name, sizeof name, "grid %dx%dx%d w%d h%d z%d", q.nx, q.ny, q.nz, q.w, q.h,
                          sh.zchunk);
            rc |= check(name, a, sh, 8);
        }
    }
    {
        // 7-point (3D) and 5-point (2D) stencils: fewer shared columns per
        // tile than the 27-point one, the same two row distances (a 2D grid
        // has one: its tiles are consecutive rows walked along it)
        struct { int nx, ny, nz, w, h; } g[] = {{20, 20, 20, 4, 8}, {20, 20, 20, 8, 4}, {13, 7, 9, 8, 8},
                                               {120, 120, 1, 8, 4}, {33, 50, 1, 4, 8}};
        for (const auto &q : g) {
            Csr a = stencil7(q.nx, q.ny, q.nz);
            RingShape sh;
            sh.w = q.w;
            sh.h = q.h;
            sh.s_mid = q.nx;
            sh.s_big = q.nz > 1 ? q.nx * q.ny : q.nx;
            if (sh.s_big > 1 && (sh.s_mid < sh.w || sh.s_mid == sh.s_big)) {
                sh.s_mid = sh.s_big;         // one distance, as the library shapes it
                sh.w = q.w * q.h;
                sh.h = 1;
                if (sh.w > sh.s_big) sh.s_big = 1;
            }
            sh.zchunk = dspmm::ring_auto_chunk(a.nr, sh, 256);
            char name[96];
            std::snprintf(name, sizeof name, "7-point %dx%dx%d w%d h%d", q.nx, q.ny, q.nz, q.w, q.h);
            rc |= check(name, a, sh, 8);
        }
    }
    {
        // no structure: one row; nine rows (a partial tile); most rows
        // empty; many more columns than rows; a few columns shared by many
        // rows (one B row read by every tile)
        struct { int nr, nc, maxlen; uint64_t seed; int w, h; const char *name; } r[] = {
            {1, 50, 20, 3, 8, 4, "one row"},
            {9, 9, 9, 5, 4, 8, "nine rows"},
            {4000, 100000, 30, 7, 8, 4, "wide: nc = 25 nr"},
            {6000, 12, 12, 9, 4, 8, "narrow: 12 columns"},
        };
        for (const auto &q : r) {
            Csr a = random_rows(q.nr, q.nc, q.maxlen, q.seed);
            RingShape sh;
            sh.w = q.w;
            sh.h = q.h;
            sh.zchunk = dspmm::ring_auto_chunk(a.nr, sh, 256);
            rc |= check(q.name, a, sh, 8);
        }
        Csr e = random_rows(3000, 3000, 6, 13);      // keep one row in 100
        Csr s;
        s.nr = e.nr;
        s.nc = e.nc;
        s.rp.push_back(0);
        for (int i = 0; i < e.nr; i++) {
            if (i % 100 == 17)
                for (int q = e.rp[(size_t)i]; q < e.rp[(size_t)i + 1]; q++) {
                    s.ci.push_back(e.ci[(size_t)q]);
                    s.va.push_back(e.va[(size_t)q]);
                }
            s.rp.push_back((int)s.ci.size());
        }
        RingShape sh;
        sh.w = 8;
        sh.h = 4;
        sh.zchunk = 4;
        rc |= check("mostly empty rows", s, sh, 8);
    }
    {
        // the library takes its row distances from sampled rows, so an
        // unstructured matrix may come with any pair: each (w, h, s_mid,
        // s_big) as the library turns it into a shape (one distance when the
        // smaller is under w) must give an exact plan, blocked or not
        Csr a = random_rows(2500, 2600, 24, 21);
        const int dist[][2] = {{8, 64}, {16, 48}, {5, 37}, {40, 40}, {1, 300}, {600, 120},
                               {4, 2400}, {2, 2}};
        for (const auto &d : dist)
            for (int w : {4, 8})
                for (int h : {4, 8}) {
                    RingShape sh;
                    sh.w = w;
                    sh.h = h;
                    sh.s_mid = d[0];
                    sh.s_big = d[1];
                    if (sh.s_big > 1 && (sh.s_mid < w || sh.s_mid == 1)) {
                        sh.s_mid = sh.s_big;
                        sh.w = w * h;
                        sh.h = 1;
                        if (sh.w > sh.s_big) sh.s_big = 1;
                    }
                    sh.zchunk = dspmm::ring_auto_chunk(a.nr, sh, 256);
                    char name[96];
                    std::snprintf(name, sizeof name, "random, distances %d %d, w%d h%d", d[0], d[1],
                                  w, h);
                    rc |= check(name, a, sh, 4);
                }
    }
    {
        // tiles per stream for a GPU's CU count, worked by hand: a 100^3
        // grid in 4 x 8 blocks is 25 x 13 = 325 walks of 100 planes
        RingShape sh;
        sh.w = 4;
        sh.h = 8;
        sh.s_mid = 100;
        sh.s_big = 10000;
        struct { int nr, cus, want; const char *why; } cases[] = {
            // 34: 975 walks, ceil(975 / 256) = 4 per CU, 4 x (2 + 34) = 144;
            // 17: 8 x 19 = 152; 15: 9 x 17 = 153; 10: 13 x 12 = 156; 50: 3 x 52
            {1000000, 256, 34, "325 walks, 256 CUs"},
            // 34: 2 x 36 = 72; 17: 4 x 19 = 76; 25: 3 x 27 = 81; 100: 1 x 102
            {1000000, 512, 34, "325 walks, 512 CUs"},
            // every split fits one walk per CU: the shortest walks
            {1000000, 100000, 1, "one walk per CU at any length"},
        };
        for (const auto &c : cases) {
            const int got = dspmm::ring_auto_chunk(c.nr, sh, c.cus);
            std::printf("auto chunk (%s): %d (want %d)\n", c.why, got, c.want);
            if (got != c.want) rc |= 1;
        }
        // consecutive tiles: 1000 rows in 32-row tiles = 32 tiles, 4 CUs:
        // 8 per walk is 4 walks, one per CU, cost 10; 7 makes 5 walks (2 on
        // a CU, 18); 16 costs 18
        RingShape c1;
        c1.w = 8;
        c1.h = 4;
        const int got = dspmm::ring_auto_chunk(1000, c1, 4);
        std::printf("auto chunk (32 consecutive tiles, 4 CUs): %d (want 8)\n", got);
        if (got != 8) rc |= 1;
    }
    return rc;
}
