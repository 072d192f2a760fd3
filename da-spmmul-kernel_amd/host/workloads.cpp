// workloads.cpp -- seeded synthetic inputs (see workloads.h).
//
// RNG: SplitMix64 (Steele, Lea, Flood 2014) expands a 64-bit key into the
// state of xoshiro256** (Blackman & Vigna 2018).  Normals use Box-Muller;
// Zipf and small-mean binomials use inversion of the exact CDF; larger
// binomials use the normal approximation (stated where it is used).
#include "workloads.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t a, uint64_t b) {
        uint64_t x = seed;
        x ^= splitmix64(a) + 0x632BE59BD9B4E019ull;
        x ^= rotl(splitmix64(b), 17);
        for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
    }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }          // [0,1)
    float uf(float lo, float hi) {                                        // [lo,hi)
        const float u = (float)(next() >> 40) * 0x1.0p-24f;
        return lo + (hi - lo) * u;
    }
    uint64_t below(uint64_t n) {                                          // [0,n)
        // Lemire's multiply-shift with rejection: unbiased.
        __uint128_t m = (__uint128_t)next() * n;
        uint64_t l = (uint64_t)m;
        if (l < n) {
            const uint64_t t = (0 - n) % n;
            while (l < t) { m = (__uint128_t)next() * n; l = (uint64_t)m; }
        }
        return (uint64_t)(m >> 64);
    }
    double normal() {
        double u1 = 1.0 - u01();          // (0,1]
        double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

enum { STREAM_LEN = 1, STREAM_COL = 2, STREAM_VAL = 3 };

struct ZipfTable {
    int kmin = 1, kmax = 1;
    std::vector<double> cdf;
    ZipfTable(double s, int lo, int hi) : kmin(lo), kmax(hi) {
        cdf.resize((size_t)(hi - lo + 1));
        double acc = 0;
        for (int k = lo; k <= hi; k++) { acc += std::pow((double)k, -s); cdf[k - lo] = acc; }
        for (auto &c : cdf) c /= acc;
    }
    int draw(Rng &r) const {
        const double u = r.u01();
        const size_t i = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        return kmin + (int)std::min(i, cdf.size() - 1);
    }
};

int binomial_draw(Rng &r, int n, double p) {
    if (p <= 0) return 0;
    if (p >= 1) return n;
    const double mean = n * p;
    if (mean < 30.0) {
        // Inversion over the exact pmf: P(k+1) = P(k) * (n-k)/(k+1) * p/(1-p).
        double u = r.u01();
        double pk = std::pow(1.0 - p, n);
        const double ratio = p / (1.0 - p);
        int k = 0;
        while (u >= pk && k < n) { u -= pk; pk *= (double)(n - k) / (k + 1) * ratio; k++; }
        return k;
    }
    // Normal approximation for mean >= 30 (error well below what the
    // benchmark's row-length distribution cares about).
    const double x = std::floor(mean + std::sqrt(mean * (1.0 - p)) * r.normal() + 0.5);
    return (int)std::min<double>(n, std::max(0.0, x));
}

int row_length(const dspmm_gen_params &p, const ZipfTable *zt, int row) {
    switch (p.kind) {
    case DSPMM_GEN_FIXED:
        return std::min(p.nc, (int)std::llround(p.p0));
    case DSPMM_GEN_ZIPF: {
        Rng r(p.seed, (uint64_t)row, STREAM_LEN);
        return std::min(p.nc, zt->draw(r));
    }
    case DSPMM_GEN_STENCIL27: {
        const int n = (int)p.p0;
        const int x = row % n, y = (row / n) % n, z = row / (n * n);
        auto span = [n](int c) { return (c > 0 ? 1 : 0) + 1 + (c < n - 1 ? 1 : 0); };
        return span(x) * span(y) * span(z);
    }
    case DSPMM_GEN_BINOMIAL: {
        Rng r(p.seed, (uint64_t)row, STREAM_LEN);
        const double d = std::min<double>(p.p0, p.nc);
        return std::max(1, binomial_draw(r, p.nc, d / p.nc));
    }
    }
    return 0;
}

// `len` distinct uniform columns in [0, nc), sorted ascending.
void draw_columns(Rng &r, int nc, int len, int *out, std::vector<int> &scratch) {
    if (len <= 0) return;
    if ((int64_t)len * 4 >= nc) {
        // Dense-ish row: selection sampling (Knuth 3.4.2 S) keeps order.
        int need = len;
        for (int c = 0; c < nc && need > 0; c++) {
            if (r.below((uint64_t)(nc - c)) < (uint64_t)need) { *out++ = c; need--; }
        }
        return;
    }
    scratch.clear();
    while ((int)scratch.size() < len) {
        const int want = len - (int)scratch.size();
        for (int i = 0; i < want; i++) scratch.push_back((int)r.below((uint64_t)nc));
        std::sort(scratch.begin(), scratch.end());
        scratch.erase(std::unique(scratch.begin(), scratch.end()), scratch.end());
    }
    std::memcpy(out, scratch.data(), sizeof(int) * (size_t)len);
}

}  // namespace

extern "C" {

int64_t dspmm_gen_rowptr(const dspmm_gen_params *p, int *rowPtr) {
    if (!p || p->nr < 0 || p->nc <= 0) return -1;
    if (p->kind == DSPMM_GEN_STENCIL27) {
        const int64_t n = (int64_t)p->p0;
        if (n * n * n != p->nr || p->nr != p->nc) return -1;
    }
    ZipfTable *zt = nullptr;
    if (p->kind == DSPMM_GEN_ZIPF) zt = new ZipfTable(p->p0, (int)p->p1, (int)p->p2);
    std::vector<int> len((size_t)p->nr);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < p->nr; r++) len[(size_t)r] = row_length(*p, zt, r);
    delete zt;
    int64_t acc = 0;
    for (int r = 0; r < p->nr; r++) {
        if (rowPtr) rowPtr[r] = (int)std::min<int64_t>(acc, INT32_MAX);
        acc += len[(size_t)r];
    }
    if (acc > INT32_MAX) return -1;
    if (rowPtr) rowPtr[p->nr] = (int)acc;
    return acc;
}

int dspmm_gen_fill(const dspmm_gen_params *p, const int *rowPtr, int *colIdx,
                   float *values) {
    if (!p || !rowPtr) return 1;
    const int n = (int)p->p0;
#pragma omp parallel
    {
        std::vector<int> scratch;
#pragma omp for schedule(dynamic, 256)
        for (int r = 0; r < p->nr; r++) {
            const int b = rowPtr[r], e = rowPtr[r + 1];
            if (p->kind == DSPMM_GEN_STENCIL27) {
                const int x = r % n, y = (r / n) % n, z = r / (n * n);
                int q = b;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int xx = x + dx, yy = y + dy, zz = z + dz;
                            if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
                            colIdx[q++] = xx + n * (yy + n * zz);
                        }
            } else {
                Rng rc(p->seed, (uint64_t)r, STREAM_COL);
                draw_columns(rc, p->nc, e - b, colIdx + b, scratch);
            }
            Rng rv(p->seed, (uint64_t)r, STREAM_VAL);
            for (int q = b; q < e; q++)
                values[q] = (p->value_dist == DSPMM_VAL_NORMAL) ? (float)rv.normal()
                                                                : rv.uf(p->lo, p->hi);
        }
    }
    return 0;
}

void dspmm_gen_rowidx(int nr, const int *rowPtr, int *rowIdx) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nr; r++)
        for (int q = rowPtr[r]; q < rowPtr[r + 1]; q++) rowIdx[q] = r;
}

void dspmm_gen_dense_uniform(uint64_t seed, int64_t n, float lo, float hi, float *out) {
    const int64_t chunk = 1 << 16;
    const int64_t nchunks = (n + chunk - 1) / chunk;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nchunks; c++) {
        Rng r(seed, (uint64_t)c, 0x44454E5345ull);
        const int64_t e = std::min(n, (c + 1) * chunk);
        for (int64_t i = c * chunk; i < e; i++) out[i] = r.uf(lo, hi);
    }
}

void dspmm_gen_batch_params(uint64_t seed, int count, int mlo, int mhi, double mu,
                            double sigma, double dlo, double dhi, const int *ncand,
                            int nncand, int *m, double *d, int *n) {
    Rng r(seed, 0xBA7C4ull, 0);
    for (int i = 0; i < count; i++) {
        m[i] = mlo + (int)r.below((uint64_t)(mhi - mlo + 1));
        const double x = std::exp(mu + sigma * r.normal());
        d[i] = std::min(std::max(x, dlo), std::min(dhi, (double)m[i]));
        n[i] = ncand[r.below((uint64_t)nncand)];
    }
}

void dspmm_gen_arrivals(uint64_t seed, int count, double rate, double p_large,
                        double *t, int *cls) {
    Rng r(seed, 0xA771ull, 0);
    double now = 0;
    for (int i = 0; i < count; i++) {
        now += -std::log(1.0 - r.u01()) / rate;    // exponential gaps
        t[i] = now;
        cls[i] = (r.u01() < p_large) ? 1 : 0;
    }
}

double dspmm_gen_zipf_mean(double s, int kmin, int kmax) {
    double num = 0, den = 0;
    for (int k = kmin; k <= kmax; k++) {
        const double w = std::pow((double)k, -s);
        num += k * w;
        den += w;
    }
    return num / den;
}

}  // extern "C"
