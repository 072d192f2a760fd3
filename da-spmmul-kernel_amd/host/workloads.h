/*
 * workloads.h -- seeded synthetic sparse/dense inputs for the SpMM benchmark
 * configurations (host C++, built into libdspmm_gen.so).
 *
 * The reference loads one Matrix-Market file per run and fills the dense
 * operand from an unseeded rand() (helper.h:66-76, :107-116).  These
 * generators replace that with reproducible inputs: every row's length,
 * columns and values come from an xoshiro256** stream keyed by
 * (seed, row, purpose), so the output does not depend on thread count.
 *
 * Row patterns (all produce sorted, duplicate-free column lists):
 *   DSPMM_GEN_FIXED     exactly round(p0) distinct uniform columns per row
 *   DSPMM_GEN_ZIPF      row length ~ Zipf(s=p0) on [p1, p2], uniform columns
 *   DSPMM_GEN_STENCIL27 27-point stencil on an n^3 grid (n=p0), x fastest
 *   DSPMM_GEN_BINOMIAL  row length ~ Binomial(nc, p0/nc) clipped to >= 1
 * Values: DSPMM_VAL_UNIFORM in [lo, hi) or DSPMM_VAL_NORMAL N(0,1).
 */
#ifndef DSPMM_WORKLOADS_H
#define DSPMM_WORKLOADS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DSPMM_GEN_FIXED = 0,
    DSPMM_GEN_ZIPF = 1,
    DSPMM_GEN_STENCIL27 = 2,
    DSPMM_GEN_BINOMIAL = 3
};
enum { DSPMM_VAL_UNIFORM = 0, DSPMM_VAL_NORMAL = 1 };

typedef struct {
    int kind;
    int nr;
    int nc;
    int value_dist;
    uint64_t seed;
    double p0, p1, p2;
    float lo, hi;
} dspmm_gen_params;

/* Row lengths.  Returns nnz (int64) and writes rowPtr[nr+1] when rowPtr is
 * not NULL.  Returns -1 when nnz would not fit in int32 or on bad input. */
int64_t dspmm_gen_rowptr(const dspmm_gen_params *p, int *rowPtr);

/* Columns and values for a rowPtr made by dspmm_gen_rowptr with the same
 * parameters.  Returns 0 on success. */
int dspmm_gen_fill(const dspmm_gen_params *p, const int *rowPtr,
                   int *colIdx, float *values);

/* COO row array from rowPtr (row of every nonzero). */
void dspmm_gen_rowidx(int nr, const int *rowPtr, int *rowIdx);

/* n uniform floats in [lo, hi) (24-bit mantissa draws, exact in FP32). */
void dspmm_gen_dense_uniform(uint64_t seed, int64_t n, float lo, float hi,
                             float *out);

/* Per-matrix parameters of the batched workload: m[i] uniform on
 * [mlo, mhi], mean row length d[i] ~ LogNormal(mu, sigma) clipped to
 * [dlo, min(dhi, m[i])], n[i] drawn uniformly from ncand[0..nncand). */
void dspmm_gen_batch_params(uint64_t seed, int count, int mlo, int mhi,
                            double mu, double sigma, double dlo, double dhi,
                            const int *ncand, int nncand,
                            int *m, double *d, int *n);

/* Poisson arrival times (seconds) of `count` requests at `rate` per second,
 * and a class draw per request: cls[i] = 1 with probability p_large. */
void dspmm_gen_arrivals(uint64_t seed, int count, double rate, double p_large,
                        double *t, int *cls);

/* Mean row length of Zipf(s) clipped to [kmin, kmax] (for choosing s). */
double dspmm_gen_zipf_mean(double s, int kmin, int kmax);

#ifdef __cplusplus
}
#endif
#endif
