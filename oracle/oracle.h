/*
 * oracle.h -- CPU restatement of the reference's host-side SpMV/SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the library links, loads or calls
 * this code; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker and never as the
 * thing that is measured or shipped.
 *
 * Each function names the reference file:line it restates.  How every
 * function is pinned (closed forms, FP64 brute force, scipy golden vectors,
 * integer exactness) is listed in tests/test_oracle.py.
 */
#ifndef DSPMM_ORACLE_H
#define DSPMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dense layouts, same numeric values as the reference (helper.h:261-264). */
#define ORACLE_ROW_MAJOR 0
#define ORACLE_COL_MAJOR 1

/*
 * Host FP32 loop: restates cpuSpmvReference (spmvspmm_benchmark.cu:11-39).
 * Row-major: out[r*nv+k] = sum_p values[p]*input[colIdx[p]*nv+k]
 * Col-major: out[k*nr+r] = sum_p values[p]*input[k*nc+colIdx[p]]
 * Every output is accumulated in float, in p order, starting from 0, with a
 * separately rounded multiply and add (the file is built with
 * -ffp-contract=off, matching an x86-64 build without FMA).
 */
void oracle_spmm_host32(int layout, int nr, int nc, int nnz, int nv,
                        const int *rowPtr, const int *colIdx,
                        const float *values, const float *input, float *out);

/*
 * FP64 oracle of the same product (no reference counterpart; the north
 * star's error reference).  out64 gets the double-precision result and,
 * when absSum is not NULL, absSum gets sum_p |values[p]*input[..]| per
 * output (the condition scale used by the relative-error gate).  Both use
 * the layout's packing (ld=nv row-major, ld=nr col-major for outputs).
 */
void oracle_spmm_fp64(int layout, int nr, int nc, int nnz, int nv,
                      const int *rowPtr, const int *colIdx,
                      const float *values, const float *input,
                      double *out64, double *absSum);

/*
 * Rounds n floats to bfloat16 (round-to-nearest-even on the FP32 bit
 * pattern, NaN kept quiet) and returns them widened back to float.  Lets a
 * test separate the rounding of a BF16-stored dense operand from the
 * accumulation error of the kernel.
 */
void oracle_round_bf16(int64_t n, const float *in, float *out);

/*
 * COO row array -> CSR rowPtr[nr+1]: restates generateCsrRowPtrArray
 * (spmvspmm_benchmark.cu:41-76), including its run-length walk over the
 * sorted row array, its zero-filled leading rows and its trailing fill.
 * Requires nnz >= 1 and sorted rows in [0,nr) as the reference does
 * (it reads rowIdxBuf[0] unconditionally, :44).  Returns rowPtr[nr].
 */
int oracle_csr_rowptr(int nr, int nnz, const int *rowIdx, int *rowPtr);

/*
 * Absolute-tolerance comparison: restates checkResult<host,dev>
 * (spmvspmm_benchmark.cu:93-116).  ref is laid out as host_layout, out as
 * dev_layout, both nrows x ncols.  Returns 1 on pass.  On the first
 * failure (the reference stops there, :107-108) *fi,*fj get its index.
 */
int oracle_check_result(int host_layout, int dev_layout, int nrows, int ncols,
                        const float *ref, const float *out, float eps,
                        int *fi, int *fj);

/*
 * Times oracle_spmm_host32 over `reps` repetitions on the calling thread
 * and returns seconds per repetition (steady clock).  Used by the CPU
 * baseline leg only.
 */
double oracle_time_host32(int layout, int nr, int nc, int nnz, int nv,
                          const int *rowPtr, const int *colIdx,
                          const float *values, const float *input,
                          float *out, int reps);

#ifdef __cplusplus
}
#endif
#endif
