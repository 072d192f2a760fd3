/* dspmm_rocsparse.h -- the vendor-library comparator beside the hand-written
 * kernels: rocSPARSE's generic SpMV/SpMM on the same device arrays.
 *
 * Replaces CusparseSpmmProblem (cusparse_helper.h:120-437): its constructor
 * (:143-178), run(format, layout, nv, alg) (:215-231, :360-384) and
 * benchmark(...) (:233-271, :385-449), as driven by the sweep's
 * cusparse_* columns (spmvspmm_benchmark.cu:304-322, run.sh:8).  Built as its
 * own library (lib/libdspmm_rocsparse.so) so the kernel library does not
 * depend on rocSPARSE.
 *
 * A "column" is one of the reference's comparator columns.  For nv == 1
 * there are 4 (SpMV), otherwise 10 (SpMM):
 *
 *   nv == 1  0 csr  row-major  default           (CUSPARSE_MV_ALG_DEFAULT)
 *            1 csr  row-major  csr_adaptive      (CUSPARSE_CSRMV_ALG1)
 *            2 coo  row-major  default           (CUSPARSE_MV_ALG_DEFAULT)
 *            3 coo  row-major  coo_atomic        (CUSPARSE_COOMV_ALG)
 *   nv > 1   0 csr  row-major  default           (CUSPARSE_SPMM_ALG_DEFAULT)
 *            1 csr  row-major  csr_row_split     (CUSPARSE_SPMM_CSR_ALG2)
 *            2 coo  row-major  default           (CUSPARSE_SPMM_ALG_DEFAULT)
 *            3 coo  row-major  coo_segmented_atomic (CUSPARSE_SPMM_COO_ALG4)
 *            4 csr  col-major  default           (CUSPARSE_SPMM_ALG_DEFAULT)
 *            5 csr  col-major  csr               (CUSPARSE_SPMM_CSR_ALG1)
 *            6 coo  col-major  default           (CUSPARSE_SPMM_ALG_DEFAULT)
 *            7 coo  col-major  coo_atomic        (CUSPARSE_SPMM_COO_ALG1)
 *            8 coo  col-major  coo_segmented     (CUSPARSE_SPMM_COO_ALG2)
 *            9 coo  col-major  coo_segmented_atomic (CUSPARSE_SPMM_COO_ALG3)
 *
 * Dense operands follow the reference: row-major B is nc x nv with ld = nv,
 * column-major B has ld = nc; C likewise with nr.  C = A * B (beta = 0).
 * A combination rocSPARSE rejects returns DSPMM_STATUS_NOT_SUPPORTED (the
 * reference exits the process instead, cusparse_helper.h:4-14).
 */
#ifndef DSPMM_ROCSPARSE_H
#define DSPMM_ROCSPARSE_H

#include "dspmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dspmm_rocsparse_s *dspmm_rocsparse_t;

/* All array arguments are device pointers and stay owned by the caller.
 * rowIdx may be NULL (the COO columns then return NOT_SUPPORTED). */
DSPMM_API dspmm_status_t dspmm_rocsparse_create(dspmm_rocsparse_t *h, int nr, int nc, int nnz,
                                                const int *rowPtr, const int *rowIdx,
                                                const int *colIdx, const float *values,
                                                const float *dnInput, float *dnOutput,
                                                dspmm_stream_t stream);
DSPMM_API dspmm_status_t dspmm_rocsparse_destroy(dspmm_rocsparse_t h);
/* Number of comparator columns for this nv (4 or 10), and their names
 * ("csr_row_default", ...). */
DSPMM_API int dspmm_rocsparse_columns(int nv);
DSPMM_API const char *dspmm_rocsparse_column_name(int nv, int column);
/* The column's layout (DENSE_ROW_MAJOR / DENSE_COL_MAJOR). */
DSPMM_API int dspmm_rocsparse_column_layout(int nv, int column);
/* One C = A * B on the handle's stream (asynchronous). */
DSPMM_API dspmm_status_t dspmm_rocsparse_run(dspmm_rocsparse_t h, int nv, int column);
/* Mean milliseconds per call over repeatIters calls after warmupIters. */
DSPMM_API dspmm_status_t dspmm_rocsparse_benchmark(dspmm_rocsparse_t h, int nv, int column,
                                                   int warmupIters, int repeatIters,
                                                   float *ms);

#ifdef __cplusplus
}
#endif
#endif /* DSPMM_ROCSPARSE_H */
