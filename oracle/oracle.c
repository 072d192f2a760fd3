/*
 * oracle.c -- CPU restatement of the reference's host path (test
 * infrastructure only; see oracle.h).  Built with -O3 -ffp-contract=off so
 * that `acc += a*b` is a rounded multiply followed by a rounded add, like the
 * reference's host loop compiled for baseline x86-64 (Makefile:3 passes no
 * host arch flags, and baseline x86-64 has no FMA).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* spmvspmm_benchmark.cu:11-39 (cpuSpmvReference). */
void oracle_spmm_host32(int layout, int nr, int nc, int nnz, int nv,
                        const int *rowPtr, const int *colIdx,
                        const float *values, const float *input, float *out)
{
    (void)nnz;
    if (layout == ORACLE_COL_MAJOR) {
        /* :13-25 -- one dense column at a time, rows in order. */
        for (int k = 0; k < nv; k++) {
            const float *vec = input + (int64_t)k * nc;
            for (int r = 0; r < nr; r++) {
                float val = 0.0f;
                for (int p = rowPtr[r]; p < rowPtr[r + 1]; p++)
                    val += values[p] * vec[colIdx[p]];
                out[(int64_t)k * nr + r] = val;
            }
        }
    } else {
        /* :26-38 -- one row at a time, nv running sums per row.  The
         * reference allocates a std::vector per row (:28); a single scratch
         * array gives the same arithmetic. */
        float *vals = (float *)malloc(sizeof(float) * (size_t)(nv > 0 ? nv : 1));
        for (int r = 0; r < nr; r++) {
            for (int k = 0; k < nv; k++) vals[k] = 0.0f;
            for (int p = rowPtr[r]; p < rowPtr[r + 1]; p++) {
                const float a = values[p];
                const float *brow = input + (int64_t)colIdx[p] * nv;
                for (int k = 0; k < nv; k++)
                    vals[k] += a * brow[k];
            }
            for (int k = 0; k < nv; k++)
                out[(int64_t)r * nv + k] = vals[k];
        }
        free(vals);
    }
}

/* FP64 oracle: same sums in double, plus sum |a*b| per output. */
void oracle_spmm_fp64(int layout, int nr, int nc, int nnz, int nv,
                      const int *rowPtr, const int *colIdx,
                      const float *values, const float *input,
                      double *out64, double *absSum)
{
    (void)nnz;
    double *acc = (double *)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    double *aacc = (double *)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    for (int r = 0; r < nr; r++) {
        for (int k = 0; k < nv; k++) { acc[k] = 0.0; aacc[k] = 0.0; }
        for (int p = rowPtr[r]; p < rowPtr[r + 1]; p++) {
            const double a = (double)values[p];
            const int c = colIdx[p];
            for (int k = 0; k < nv; k++) {
                const double b = (layout == ORACLE_COL_MAJOR)
                                     ? (double)input[(int64_t)k * nc + c]
                                     : (double)input[(int64_t)c * nv + k];
                acc[k] += a * b;
                aacc[k] += fabs(a * b);
            }
        }
        for (int k = 0; k < nv; k++) {
            const int64_t o = (layout == ORACLE_COL_MAJOR)
                                  ? (int64_t)k * nr + r
                                  : (int64_t)r * nv + k;
            out64[o] = acc[k];
            if (absSum) absSum[o] = aacc[k];
        }
    }
    free(acc);
    free(aacc);
}

void oracle_round_bf16(int64_t n, const float *in, float *out)
{
    for (int64_t i = 0; i < n; i++) {
        uint32_t u;
        memcpy(&u, &in[i], 4);
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
            u = (u | 0x00400000u) & 0xffff0000u;          /* quiet NaN */
        } else {
            const uint32_t lsb = (u >> 16) & 1u;           /* ties to even */
            u = (u + 0x7fffu + lsb) & 0xffff0000u;
        }
        memcpy(&out[i], &u, 4);
    }
}

/* spmvspmm_benchmark.cu:41-76 (generateCsrRowPtrArray): walk the sorted row
 * array run by run; after each run write the run's end offset for the run's
 * row and for every empty row up to the next run's row. */
int oracle_csr_rowptr(int nr, int nnz, const int *rowIdx, int *rowPtr)
{
    for (int i = 0; i <= nr; i++) rowPtr[i] = 0;     /* :43 */
    if (nnz <= 0) return 0;                          /* reference requires nnz>=1 */
    int curr_row = rowIdx[0];                        /* :44 */
    int curr_pos = 0;                                /* :45 */
    int it = curr_row;                               /* :47, rows <= curr_row stay 0 */
    int next_row = curr_row;   /* the reference leaves it unset until read; any
                                  value gives the same output (see tests) */
    for (; it != nr + 1;) {                          /* :63 */
        do {                                         /* :64-66 */
            ++curr_pos;
        } while (curr_pos < nnz && ((next_row = rowIdx[curr_pos]) == curr_row));
        do {                                         /* :67-70 */
            rowPtr[++it] = curr_pos;
            curr_row++;
        } while (curr_row < nr && curr_row < next_row);
        if (curr_pos == nnz) break;                  /* :71 */
    }
    while (++it < nr + 1) rowPtr[it] = curr_pos;     /* :73 */
    return rowPtr[nr];
}

/* spmvspmm_benchmark.cu:93-116 (checkResult). */
int oracle_check_result(int host_layout, int dev_layout, int nrows, int ncols,
                        const float *ref, const float *out, float eps,
                        int *fi, int *fj)
{
    for (int i = 0; i < nrows; i++) {
        for (int j = 0; j < ncols; j++) {
            const int64_t i1 = (host_layout == ORACLE_ROW_MAJOR)
                                   ? (int64_t)i * ncols + j : (int64_t)j * nrows + i;
            const int64_t i2 = (dev_layout == ORACLE_ROW_MAJOR)
                                   ? (int64_t)i * ncols + j : (int64_t)j * nrows + i;
            /* :103 -- a NaN difference compares false and passes, as in the
             * reference; tests that need finiteness check it separately. */
            if (fabs((double)(out[i2] - ref[i1])) > (double)eps) {
                if (fi) *fi = i;
                if (fj) *fj = j;
                return 0;
            }
        }
    }
    return 1;
}

double oracle_time_host32(int layout, int nr, int nc, int nnz, int nv,
                          const int *rowPtr, const int *colIdx,
                          const float *values, const float *input,
                          float *out, int reps)
{
    struct timespec t0, t1;
    if (reps < 1) reps = 1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < reps; i++)
        oracle_spmm_host32(layout, nr, nc, nnz, nv, rowPtr, colIdx, values, input, out);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec)) / reps;
}
