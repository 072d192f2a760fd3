/*
 * dspmm.h -- C-ABI of the MI355X (gfx950) SpMV/SpMM kernel library.
 *
 * Drop-in boundary for the reference's kernel family (spmv.h, spmm.h) and
 * its dispatcher (spmvspmm.h).  Every entry point takes plain pointers and
 * sizes; streams are passed as `dspmm_stream_t`, which is a hipStream_t
 * (NULL = the default stream).  No entry point exits the process: errors
 * come back as dspmm_status_t.
 *
 * Data contract (same as the reference, spmvspmm.h:13-50):
 *   - 0-based int32 indices; COO rows sorted; rowPtr[nr+1] consistent;
 *   - rowIdx (COO) is read only by the COO (nnz-balanced) algorithms;
 *   - row-major dense operands are packed with ld = nv, column-major ones
 *     with ld = nc (input) and ld = nr (output);
 *   - all pointers are caller-owned device memory; launchers allocate
 *     nothing and are asynchronous on `stream`.
 * Output semantics (spmvspmm_benchmark.cu:141-157):
 *   - CSR (row-balanced) algorithms OVERWRITE every output entry;
 *   - COO (nnz-balanced) algorithms ACCUMULATE into the output with atomic
 *     adds, so the caller zeroes it first, as the reference's harness does.
 *
 * Every function declared with DSPMM_API is exported by libdspmm.so.
 */
#ifndef DSPMM_H
#define DSPMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSPMM_API __attribute__((visibility("default")))

/* A hipStream_t; NULL selects the default stream. */
typedef void *dspmm_stream_t;

/* ---------------------------------------------------------------------
 * Enums with the reference's names and numeric values.
 * ------------------------------------------------------------------- */

/* spmvspmm.h:5-10 */
typedef enum SPMV_SPMM_ALG {
    ALG_CSR_SCALAR = 0, /* row-balanced, sequential reduction   (SR_RB) */
    ALG_CSR_VECTOR = 1, /* row-balanced, parallel reduction     (PR_RB) */
    ALG_COO_SCALAR = 2, /* nnz-balanced, sequential reduction   (SR_EB) */
    ALG_COO_VECTOR = 3  /* nnz-balanced, parallel reduction     (PR_EB) */
} SPMV_SPMM_ALG;

/* helper.h:256-259 */
typedef enum SparseFormat { SPARSE_FORMAT_CSR = 0, SPARSE_FORMAT_COO = 1 } SparseFormat;

/* helper.h:261-264 */
typedef enum DenseLayout { DENSE_ROW_MAJOR = 0, DENSE_COL_MAJOR = 1 } DenseLayout;

typedef enum dspmm_status {
    DSPMM_STATUS_SUCCESS = 0,
    DSPMM_STATUS_INVALID_VALUE = 1,     /* bad size, pointer or enum       */
    DSPMM_STATUS_NOT_SUPPORTED = 2,     /* e.g. nv outside a kernel's range */
    DSPMM_STATUS_ALLOC_FAILED = 3,      /* device or host allocation failed */
    DSPMM_STATUS_HIP_ERROR = 4,         /* see dspmm_last_hip_error()       */
    DSPMM_STATUS_WORKSPACE_TOO_SMALL = 5,
    DSPMM_STATUS_ADMISSION_REJECTED = 6,/* request exceeds the HBM budget   */
    DSPMM_STATUS_NOT_PREPARED = 7       /* a plan or buffer the call needs was
                                           not built by dspmm_matrix_prepare, and
                                           the call may not build it (stream
                                           capture, dspmm_spmm_ex_noalloc)      */
} dspmm_status_t;

DSPMM_API const char *dspmm_get_error_string(dspmm_status_t status);
/* hipError_t of the most recent DSPMM_STATUS_HIP_ERROR on this thread. */
DSPMM_API int dspmm_last_hip_error(void);
DSPMM_API const char *dspmm_version(void);

/* ---------------------------------------------------------------------
 * The 12 kernel launchers.  Names are the reference kernels' names
 * without the `_kernel` suffix and template arguments; the argument list
 * is the reference's common one (spmv.h:6-18) plus a stream.  The
 * column-major CSR kernels declare (nv, nnz) swapped in the reference
 * (spmm.h:274-275, 320-321) but are called with (nnz, nv)
 * (spmvspmm.h:169,194); these launchers use the call-site order.
 *
 * Accepted nv: spmv_*_scalar: 1.  spmv_*_vector: 1, 2, 4 (row-major).
 * spmm_*_row / spmm_*_col: any nv >= 1 (the reference needs a power of
 * two; the dispatcher below keeps that rule).
 * ------------------------------------------------------------------- */
#define DSPMM_LAUNCHER_ARGS                                                  \
    int nr, int nc, int nnz, int nv, const int *rowPtr, const int *rowIdx,   \
        const int *colIdx, const float *values, const float *dnInput,        \
        float *dnOutput, dspmm_stream_t stream

/* spmv.h:6   spmv_csr_scalar_kernel  -- y = A x, one lane per row.         */
DSPMM_API dspmm_status_t spmv_csr_scalar(DSPMM_LAUNCHER_ARGS);
/* spmv.h:44  spmv_coo_scalar_kernel  -- y += A x, nnz chunk per lane.      */
DSPMM_API dspmm_status_t spmv_coo_scalar(DSPMM_LAUNCHER_ARGS);
/* spmv.h:95  spmv_csr_vector_kernel  -- lane group per row, reduced.       */
DSPMM_API dspmm_status_t spmv_csr_vector(DSPMM_LAUNCHER_ARGS);
/* spmv.h:153 spmv_coo_vector_kernel  -- wave over nnz, segmented reduce.   */
DSPMM_API dspmm_status_t spmv_coo_vector(DSPMM_LAUNCHER_ARGS);
/* spmm.h:8   spmm_csr_scalar_row_kernel                                     */
DSPMM_API dspmm_status_t spmm_csr_scalar_row(DSPMM_LAUNCHER_ARGS);
/* spmm.h:54  spmm_csr_vector_row_kernel                                     */
DSPMM_API dspmm_status_t spmm_csr_vector_row(DSPMM_LAUNCHER_ARGS);
/* spmm.h:125 spmm_coo_scalar_row_kernel                                     */
DSPMM_API dspmm_status_t spmm_coo_scalar_row(DSPMM_LAUNCHER_ARGS);
/* spmm.h:180 spmm_coo_vector_row_kernel                                     */
DSPMM_API dspmm_status_t spmm_coo_vector_row(DSPMM_LAUNCHER_ARGS);
/* spmm.h:271 spmm_csr_scalar_col_kernel                                     */
DSPMM_API dspmm_status_t spmm_csr_scalar_col(DSPMM_LAUNCHER_ARGS);
/* spmm.h:317 spmm_csr_vector_col_kernel                                     */
DSPMM_API dspmm_status_t spmm_csr_vector_col(DSPMM_LAUNCHER_ARGS);
/* spmm.h:371 spmm_coo_scalar_col_kernel                                     */
DSPMM_API dspmm_status_t spmm_coo_scalar_col(DSPMM_LAUNCHER_ARGS);
/* spmm.h:440 spmm_coo_vector_col_kernel                                     */
DSPMM_API dspmm_status_t spmm_coo_vector_col(DSPMM_LAUNCHER_ARGS);

/* ---------------------------------------------------------------------
 * Dispatcher: replaces cudaSpmm<256,32> (spmvspmm.h:53-199) with the same
 * parameter list plus a stream.  Maps (alg, layout, nv) to a launcher the
 * way the reference does (spmv_* for nv==1, and for nv in {2,4} with the
 * row-major vector algorithms).  Keeps the reference callers' rule that
 * nv is a power of two (spmvspmm.h:209,223) but reports
 * DSPMM_STATUS_NOT_SUPPORTED instead of printing and returning.
 * ------------------------------------------------------------------- */
DSPMM_API dspmm_status_t dspmm_spmm(SPMV_SPMM_ALG alg, DenseLayout layout,
                                    int nr, int nc, int nnz, int nv,
                                    const int *rowPtr, const int *rowIdx,
                                    const int *colIdx, const float *values,
                                    const float *dnInput, float *dnOutput,
                                    dspmm_stream_t stream);

/* ---------------------------------------------------------------------
 * Device memory and timing: the owning-array upload of hgpu::CUArray
 * (helper.h:78-154) and the event timer hgpu::GpuTimer (helper.h:158-191)
 * as plain C calls, so that hosts without a HIP toolchain can drive the
 * library.
 * ------------------------------------------------------------------- */
DSPMM_API dspmm_status_t dspmm_device_alloc(void **ptr, size_t bytes);
DSPMM_API dspmm_status_t dspmm_device_free(void *ptr);
/* Synchronous copies (like CUArray::sync_device / sync_host). */
DSPMM_API dspmm_status_t dspmm_memcpy_h2d(void *dst, const void *src, size_t bytes);
DSPMM_API dspmm_status_t dspmm_memcpy_d2h(void *dst, const void *src, size_t bytes);
DSPMM_API dspmm_status_t dspmm_memcpy_d2d(void *dst, const void *src, size_t bytes,
                                          dspmm_stream_t stream);
/* Zero-fill with a kernel (safe to record in a captured graph). */
DSPMM_API dspmm_status_t dspmm_fill_zero(void *ptr, size_t bytes, dspmm_stream_t stream);
DSPMM_API dspmm_status_t dspmm_stream_create(dspmm_stream_t *stream);
DSPMM_API dspmm_status_t dspmm_stream_destroy(dspmm_stream_t stream);
DSPMM_API dspmm_status_t dspmm_stream_synchronize(dspmm_stream_t stream);
DSPMM_API dspmm_status_t dspmm_device_synchronize(void);

typedef struct dspmm_device_info {
    char name[64];
    char arch[32];          /* e.g. "gfx950:sramecc+:xnack-" */
    int compute_units;
    int wave_size;
    int clock_khz;
    int l2_bytes;
    size_t lds_per_block;
    size_t total_mem;
    size_t free_mem;
} dspmm_device_info;
DSPMM_API dspmm_status_t dspmm_get_device_info(dspmm_device_info *info);

/* Event timer: each mark records an event on `stream`; intervals between
 * consecutive marks are returned in milliseconds after a sync. */
typedef struct dspmm_timer_s *dspmm_timer_t;
DSPMM_API dspmm_status_t dspmm_timer_create(dspmm_timer_t *t, int max_marks);
DSPMM_API dspmm_status_t dspmm_timer_destroy(dspmm_timer_t t);
DSPMM_API dspmm_status_t dspmm_timer_reset(dspmm_timer_t t);
DSPMM_API dspmm_status_t dspmm_timer_mark(dspmm_timer_t t, dspmm_stream_t stream);
/* Waits for the last mark; ms[i] = time between mark i and mark i+1.  */
DSPMM_API dspmm_status_t dspmm_timer_intervals(dspmm_timer_t t, float *ms, int n);

/* SpmvSpmmProblem::benchmark (spmvspmm.h:220-243) as one C call: `warmup`
 * untimed dispatcher calls, then `reps` timed ones with an event between
 * every launch.  ms_per_launch[reps] (may be NULL) gets each launch's time,
 * *ms_mean the mean over the timed region.  Output is not re-zeroed between
 * launches (as in the reference). */
DSPMM_API dspmm_status_t dspmm_benchmark_compat(SPMV_SPMM_ALG alg, DenseLayout layout,
                                                int nr, int nc, int nnz, int nv,
                                                const int *rowPtr, const int *rowIdx,
                                                const int *colIdx, const float *values,
                                                const float *dnInput, float *dnOutput,
                                                int warmup, int reps, float *ms_mean,
                                                float *ms_per_launch, dspmm_stream_t stream);

/* SpmvSpmmProblem::dry_run (spmvspmm.h:245-273): warm-ups, then one launch
 * bracketed by roctx profiler resume/pause and a named roctx range. */
DSPMM_API dspmm_status_t dspmm_dry_run_compat(SPMV_SPMM_ALG alg, DenseLayout layout,
                                              int nr, int nc, int nnz, int nv,
                                              const int *rowPtr, const int *rowIdx,
                                              const int *colIdx, const float *values,
                                              const float *dnInput, float *dnOutput,
                                              int warmup, dspmm_stream_t stream);

/* ---------------------------------------------------------------------
 * Native API (no reference counterpart): an owning device-resident CSR
 * matrix, any nv >= 1, leading dimensions, FP32 or BF16 dense input,
 * overwrite semantics, a load-balanced plan and a per-matrix autotuner.
 * ------------------------------------------------------------------- */
typedef struct dspmm_matrix_s *dspmm_matrix_t;

typedef enum dspmm_dtype { DSPMM_R_32F = 0, DSPMM_R_16BF = 1 } dspmm_dtype_t;

/* Upload flags. */
#define DSPMM_MATRIX_WITH_COO 0x1u  /* also keep a device rowIdx array */

/* Upload: replaces the four CUArray uploads of the sparse operand
 * (spmvspmm_benchmark.cu:369-376).  h_rowPtr may be NULL when h_rowIdx
 * (sorted) is given: rowPtr is then built as generateCsrRowPtrArray does
 * (spmvspmm_benchmark.cu:41-76), including for nnz == 0. */
DSPMM_API dspmm_status_t dspmm_matrix_create(dspmm_matrix_t *A, int nr, int nc, int nnz,
                                             const int *h_rowPtr, const int *h_rowIdx,
                                             const int *h_colIdx, const float *h_values,
                                             unsigned flags);
/* Same, from arrays already in device memory (copied, so the caller keeps
 * ownership of its buffers; the handle also keeps host copies, made here,
 * from which it builds its plans).  d_rowIdx may be NULL. */
DSPMM_API dspmm_status_t dspmm_matrix_create_device(dspmm_matrix_t *A, int nr, int nc,
                                                    int nnz, const int *d_rowPtr,
                                                    const int *d_rowIdx,
                                                    const int *d_colIdx,
                                                    const float *d_values, unsigned flags,
                                                    dspmm_stream_t stream);
/* Frees the handle and everything it holds, without a device-wide
 * synchronisation: the caller makes sure no launch that reads the handle
 * is still running (as for hipFree of the arrays themselves). */
DSPMM_API dspmm_status_t dspmm_matrix_destroy(dspmm_matrix_t A);

typedef struct dspmm_matrix_desc {
    int nr, nc, nnz;
    const int *rowPtr;     /* device */
    const int *rowIdx;     /* device, NULL unless uploaded with COO */
    const int *colIdx;     /* device */
    const float *values;   /* device */
    size_t hbm_bytes;      /* device bytes owned by the handle (incl. plans) */
    int max_row_len;
    double mean_row_len;
    double cv_row_len;     /* coefficient of variation of row lengths */
} dspmm_matrix_desc;
DSPMM_API dspmm_status_t dspmm_matrix_describe(dspmm_matrix_t A, dspmm_matrix_desc *d);

/* Kernel variants of the native path. */
typedef enum dspmm_variant {
    DSPMM_VARIANT_AUTO = 0,     /* cached autotune result, else heuristic */
    DSPMM_VARIANT_NNZ_TILED = 1,/* nnz-balanced row-aligned tiles, sequential
                                   per-row sums, deterministic, overwrite  */
    DSPMM_VARIANT_CSR_SCALAR = 2,/* a lane group per row, sequential sum (the
                                    host loop's order); row-major with unroll
                                    2/4/8/16 or BF16 B: each wavefront stages
                                    its rows' (col, val) in LDS and keeps
                                    `unroll` gathers in flight per group
                                    (tile_nnz 1: windows split into one slice
                                    per unfinished row, for rows that overflow
                                    a wavefront's 512-pair window; 0: windows
                                    of the wavefront's contiguous pairs);
                                    unroll 0 with FP32 B: the reference-class
                                    kernel (workgroup-staged, 4 in flight) */
    DSPMM_VARIANT_CSR_VECTOR = 3,
    DSPMM_VARIANT_COO_SCALAR = 4,
    DSPMM_VARIANT_COO_VECTOR = 5,
    DSPMM_VARIANT_SPMV_TILED = 6,/* nv == 1 only: workgroup per nnz tile, products
                                    in LDS, per-row sums (row-exact), tree for
                                    long rows; tile_nnz = tile size */
    DSPMM_VARIANT_NNZ_CACHED = 7,/* row-major, lanes * vec >= nv: tiles of whole
                                    rows whose distinct columns' B rows are
                                    staged in LDS once per tile (matrices whose
                                    neighbouring rows share columns); rows are
                                    summed like the host loop (row-exact);
                                    tile_nnz = distinct columns per tile,
                                    unroll = workgroup size (64/128/256,
                                    0 = from the plan's rows per tile) */
    DSPMM_VARIANT_ROW_BLOCKED = 8, /* row-major, sorted rows without repeated
                                    columns: a lane group owns tile_nnz = 2, 4
                                    or 8 rows, row_stride apart (consecutive
                                    by default), and loads each B row their
                                    columns share once; rows are summed like
                                    the host loop (row-exact); unroll = B rows
                                    in flight (4 or 8) */
    DSPMM_VARIANT_RING = 9       /* row-major, nv a multiple of 4 (FP32) or 8
                                    (BF16) fitting one slab: B rows staged in a
                                    ring in LDS whose fills and evictions a host
                                    plan decides; a workgroup walks tiles of
                                    tile_nnz consecutive rows x unroll runs
                                    (the matrix's sampled row distances apart),
                                    row_stride tiles per walk (0 = the length
                                    whose walks fill this GPU's CUs x
                                    workgroups per CU in the fewest whole
                                    rounds); rows are summed like the host
                                    loop (row-exact) */
} dspmm_variant_t;

typedef struct dspmm_config {
    int variant;    /* dspmm_variant_t */
    int vec;        /* floats per lane along N: 1, 2, 4 (8 for BF16); 0 = auto */
    int lanes;      /* lanes per row slab (power of two <= 64); 0 = auto */
    int tile_nnz;   /* NNZ_TILED: max nonzeros per tile; 0 = auto */
    int unroll;     /* NNZ_TILED: gathers in flight per lane group; 0 = auto */
    int row_stride; /* ROW_BLOCKED: distance between the rows of a block, in
                       [1, max(1, nr)]; 0 = consecutive rows; -1 = the
                       distance at which the matrix's sampled rows share the
                       most columns (a grid stencil's plane length; 1 when
                       no distance stands out).  The autotuner tries
                       consecutive rows and that distance.  RING: tiles per
                       walk (0 = chosen for the GPU).  Other variants ignore
                       it. */
} dspmm_config;

/* C = A * B, overwriting C (nv columns; ldb/ldc in elements, >= nv for
 * row-major, >= nc / nr for column-major).  btype selects FP32 or BF16 B
 * storage; accumulation is always FP32.  cfg may be NULL (auto).  The
 * row-major kernels (NNZ_TILED, ROW_BLOCKED, NNZ_CACHED and the
 * row-sequential CSR_SCALAR) take column-major operands through transposed
 * copies of B and C (results equal the row-major call's bit for bit); the
 * reference classes run column-major natively, FP32 only.  `workspace` may
 * be NULL: the handle's own buffers (split rows, transposed copies) are
 * then used, so calls on one handle must then not overlap in time.  Plans
 * are built on first use unless dspmm_matrix_prepare built them. */
DSPMM_API dspmm_status_t dspmm_spmm_ex(dspmm_matrix_t A, DenseLayout layout, int nv,
                                       const void *B, int64_t ldb, dspmm_dtype_t btype,
                                       float *C, int64_t ldc, const dspmm_config *cfg,
                                       void *workspace, size_t workspace_bytes,
                                       dspmm_stream_t stream);
/* dspmm_spmm_ex that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc.
 * It calls no hipMalloc / hipFree / allocating routine: its scratch (the
 * split-row carries and the transposed copies of column-major operands)
 * is a static device array of DSPMM_NOALLOC_SCRATCH_BYTES that the library
 * holds on every GPU it runs on until the process ends, and the plans and
 * AUTO's pick must already exist: call dspmm_matrix_prepare with the same
 * arguments first (else DSPMM_STATUS_NOT_PREPARED).  A call whose scratch
 * would exceed DSPMM_NOALLOC_SCRATCH_BYTES returns
 * DSPMM_STATUS_WORKSPACE_TOO_SMALL.  Every call on one GPU shares that
 * GPU's array (each call writes what it reads): calls that overlap in time
 * on one GPU race and must be serialised by the caller. */
#define DSPMM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
DSPMM_API dspmm_status_t dspmm_spmm_ex_noalloc(dspmm_matrix_t A, DenseLayout layout, int nv,
                                               const void *B, int64_t ldb, dspmm_dtype_t btype,
                                               float *C, int64_t ldc, const dspmm_config *cfg,
                                               dspmm_stream_t stream);
/* Workspace bytes dspmm_spmm_ex needs for these arguments; for AUTO, the
 * larger of the resolved configuration's and the static default's (the
 * fallback when the resolved one cannot run the operands). */
DSPMM_API dspmm_status_t dspmm_spmm_ex_workspace_size(dspmm_matrix_t A, DenseLayout layout,
                                                      int nv, dspmm_dtype_t btype,
                                                      const dspmm_config *cfg,
                                                      size_t *bytes);
/* Builds everything a later dspmm_spmm_ex call with these arguments needs
 * -- the tile plan of the configuration it resolves to and, for AUTO, the
 * plan of the static default it falls back to; the row array of the COO
 * classes -- and fixes AUTO's pick for (layout, nv, btype).  With
 * DSPMM_PREPARE_HANDLE_BUFFERS it also sizes the handle's own split-row and
 * transposed-copy buffers, for calls that pass no workspace.  Afterwards
 * such a call allocates nothing, copies nothing to the host and never
 * waits, so it may be captured into a hipGraph.  B and C are only read
 * for their alignment.  Plans and buffers are built by blocking copies:
 * call it outside stream capture.  An unprepared dspmm_spmm_ex builds what
 * it needs on first use, except under stream capture, where it returns
 * DSPMM_STATUS_NOT_PREPARED instead (no reference counterpart: the
 * reference dispatcher, spmvspmm.h:53-199, launches from raw arrays and
 * has no plans). */
#define DSPMM_PREPARE_HANDLE_BUFFERS 0x1u
DSPMM_API dspmm_status_t dspmm_matrix_prepare(dspmm_matrix_t A, DenseLayout layout, int nv,
                                              const void *B, int64_t ldb, dspmm_dtype_t btype,
                                              const float *C, int64_t ldc,
                                              const dspmm_config *cfg, unsigned flags);
/* Resolve what DSPMM_VARIANT_AUTO would run for these arguments. */
DSPMM_API dspmm_status_t dspmm_spmm_ex_resolve(dspmm_matrix_t A, DenseLayout layout, int nv,
                                               dspmm_dtype_t btype, const dspmm_config *cfg,
                                               dspmm_config *resolved);

/* Autotuner: times the candidate variants and tile shapes on the given
 * operands (C is overwritten; reps calls each, median), times the front
 * runners within 5% of the best again (up to four, 4 x reps calls, in both
 * orders), caches the fastest in the handle for (layout, nv, btype) and
 * returns it with its median time. */
DSPMM_API dspmm_status_t dspmm_autotune(dspmm_matrix_t A, DenseLayout layout, int nv,
                                        const void *B, int64_t ldb, dspmm_dtype_t btype,
                                        float *C, int64_t ldc, int reps,
                                        dspmm_config *best, float *best_ms,
                                        dspmm_stream_t stream);
/* Stores cfg as the tuned configuration for (layout, nv, btype), as
 * dspmm_autotune would (e.g. one saved from an earlier process); AUTO calls
 * and batches built afterwards use it.  cfg->variant AUTO clears it. */
DSPMM_API dspmm_status_t dspmm_matrix_set_config(dspmm_matrix_t A, DenseLayout layout, int nv,
                                                 dspmm_dtype_t btype, const dspmm_config *cfg);
/* The autotuner's whole table: every candidate it would time for these
 * operands and its median ms over `reps` calls (negative: not supported).
 * With capacity == 0 only *count is set; a capacity below *count gives
 * DSPMM_STATUS_WORKSPACE_TOO_SMALL.  Nothing is cached.  This is the data a
 * sweep records and the selector is fitted on. */
DSPMM_API dspmm_status_t dspmm_autotune_table(dspmm_matrix_t A, DenseLayout layout, int nv,
                                              const void *B, int64_t ldb, dspmm_dtype_t btype,
                                              float *C, int64_t ldc, int reps,
                                              dspmm_config *configs, float *ms, int capacity,
                                              int *count, dspmm_stream_t stream);

/* Timed loop over dspmm_spmm_ex: warmup untimed calls, then reps timed
 * ones.  *ms_mean is the timed region's length / reps.  ms_per_launch[reps]
 * (may be NULL) gets each call's time, ms_main_per_launch[reps] (may be
 * NULL) the time of its main kernel only, without the split-row combine that
 * follows it.  Events are recorded only where an output asks for them:
 * between calls when either array is given, and between a call's main
 * kernel and its combine for ms_main_per_launch; with both NULL only the
 * region is bracketed.  No work is skipped in any mode. */
DSPMM_API dspmm_status_t dspmm_benchmark_ex(dspmm_matrix_t A, DenseLayout layout, int nv,
                                            const void *B, int64_t ldb, dspmm_dtype_t btype,
                                            float *C, int64_t ldc, const dspmm_config *cfg,
                                            int warmup, int reps, float *ms_mean,
                                            float *ms_per_launch, float *ms_main_per_launch,
                                            dspmm_stream_t stream);

/* FP32 -> BF16 (round to nearest even) on the device; dst holds n uint16. */
DSPMM_API dspmm_status_t dspmm_convert_f32_to_bf16(const float *src, void *dst, int64_t n,
                                                   dspmm_stream_t stream);


/* ---------------------------------------------------------------------
 * Batched SpMM (no reference counterpart): many independent row-major
 * products C_i = A_i * B_i, each with its own matrix, nv and operands,
 * in ONE launch (plus one launch that combines split rows).  Each item
 * runs the nnz-tiled kernel with the static default vector width, lanes and
 * tile for its nv and dense type (autotuned single-launch configurations
 * are not used); dspmm_batch_item_config returns it, and dspmm_spmm_ex with
 * it gives the same bits.  Work is cut into wavefront units, slab-major
 * within an item; each XCD owns a list of them and walks it in order, and a
 * matrix is kept on one XCD unless it is heavy (then its units are cut into
 * eight contiguous ranges, one per XCD), so that a matrix whose B fits in
 * an XCD's L2 is gathered from there.  The default grid is one workgroup
 * per four units; blocks_per_cu > 0 gives a persistent grid.  The matrices
 * must outlive the batch.
 * ------------------------------------------------------------------- */
typedef struct dspmm_batch_s *dspmm_batch_t;

typedef struct dspmm_batch_item {
    dspmm_matrix_t A;
    int nv;
    const void *B;   /* row-major, ldb >= nv; FP32 or BF16 per the batch */
    int64_t ldb;
    float *C;        /* row-major, ldc >= nv; overwritten */
    int64_t ldc;
} dspmm_batch_item;

typedef struct dspmm_batch_info {
    int count;
    int units;           /* wavefront work units */
    int segments;
    int splits;          /* split rows combined after the main launch */
    int blocks;          /* workgroups of the one launch */
    int units_per_xcd[8];
    size_t table_bytes;  /* device bytes of the schedule tables */
    size_t carry_bytes;  /* device bytes of split-row partial sums */
} dspmm_batch_info;

/* blocks_per_cu > 0: a persistent grid of that many workgroups per CU that
 * walk the per-XCD work lists; 0: one workgroup per four work units (the
 * same schedule, spread over as many workgroups as it needs). */
DSPMM_API dspmm_status_t dspmm_batch_create(dspmm_batch_t *out, const dspmm_batch_item *items,
                                            int count, dspmm_dtype_t btype, int blocks_per_cu);
DSPMM_API dspmm_status_t dspmm_batch_destroy(dspmm_batch_t b);
DSPMM_API dspmm_status_t dspmm_batch_describe(dspmm_batch_t b, dspmm_batch_info *info);
/* The batch's one launch (+ split combine) on `stream`. */
DSPMM_API dspmm_status_t dspmm_batch_run(dspmm_batch_t b, dspmm_stream_t stream);
/* The configuration item `index` runs with in the batch (the nnz-tiled
 * kernel's static default); a single dspmm_spmm_ex call with it gives the
 * same bits.
 * No counterpart in the reference; for callers that check or replay one
 * item alone. */
DSPMM_API dspmm_status_t dspmm_batch_item_config(dspmm_batch_t b, int index, dspmm_config *out);
/* Replays a hipGraph of dspmm_batch_run (captured on first call). */
DSPMM_API dspmm_status_t dspmm_batch_graph_launch(dspmm_batch_t b, dspmm_stream_t stream);
/* Baseline: one dspmm_spmm_ex launch per item, same configs, in order. */
DSPMM_API dspmm_status_t dspmm_batch_run_plain(dspmm_batch_t b, dspmm_stream_t stream);
/* Rebinds item `index` to new operands (same alignment class); the schedule
 * and any captured graph stay valid.  Ordered on `stream`. */
DSPMM_API dspmm_status_t dspmm_batch_set_operands(dspmm_batch_t b, int index, const void *B,
                                                  float *C, dspmm_stream_t stream);
/* mode 0: dspmm_batch_run, 1: dspmm_batch_graph_launch, 2: dspmm_batch_run_plain
 * (one launch per item with the batch's configuration), 3: one launch per item
 * with the item's own AUTO configuration (tuned, else selected, else the
 * static default) -- the plain baseline at each matrix's best known kernel.
 * ms_per_launch[reps] (may be NULL) gets each pass's time. */
DSPMM_API dspmm_status_t dspmm_benchmark_batch(dspmm_batch_t b, int mode, int warmup, int reps,
                                               float *ms_mean, float *ms_per_launch,
                                               dspmm_stream_t stream);


/* ---------------------------------------------------------------------
 * Serving loop (no reference counterpart).  A server thread keeps CSR
 * matrices resident in HBM (an LRU cache bounded by cache_bytes), admits
 * queued requests while their dense operands fit an HBM arena of
 * operand_bytes, and runs continuous batching: whenever one of
 * max_inflight stream slots is free, every admissible queued request (up
 * to max_batch) is launched as one batched SpMM on that slot, so batches
 * overlap on the device and requests that arrive meanwhile join the next
 * one.  A matrix's part of a batch schedule is computed on its first
 * request with a given nv and kept while it stays resident, so admission
 * copies it.  Up to four admitted requests whose matrix's own AUTO kernel
 * sums each row in nonzero order (and whose batch plan splits no row)
 * launch one by one with that kernel instead, which at low load is faster
 * than the batch's static kernel and gives the same bits.  Results are
 * bit-identical to running the request alone (the kernels are
 * deterministic and each matrix keeps its config).  In plain
 * mode every request is one dspmm_spmm_ex launch on a single stream (the
 * baseline).  Times are seconds on dspmm_server_now()'s steady clock.
 * ------------------------------------------------------------------- */
typedef struct dspmm_server_s *dspmm_server_t;

typedef struct dspmm_server_config {
    size_t cache_bytes;     /* HBM budget of resident matrices (CSR + plans) */
    size_t operand_bytes;   /* HBM arena for server-owned B, C and split carries */
    int max_inflight;       /* batches (slots / streams) in flight, >= 1 */
    int max_batch;          /* requests per batch, >= 1 */
    int plain;              /* 0: continuous batching; 1: one launch per request
                               with the batch's configuration; 2: one launch per
                               request with the matrix's own AUTO configuration
                               (tuned, else selected) */
    dspmm_dtype_t btype;    /* storage of server-generated B (FP32 or BF16) */
    int64_t alone_nnz;      /* batching mode: requests on matrices with at least
                               this many nonzeros launch alone with the matrix's
                               own AUTO configuration (0: 16 Mi nonzeros) */
} dspmm_server_config;

typedef struct dspmm_request {
    uint64_t id;
    int matrix;         /* id from dspmm_server_add_matrix */
    int nv;             /* dense columns */
    uint64_t seed;      /* B is generated on the device from this seed (U[-1,1),
                           row-major, ld = nv) unless B is given */
    const void *B;      /* optional caller-owned device B (row-major, ld = nv) */
    float *C;           /* optional caller-owned device C (row-major, ld = nv);
                           NULL: server-owned, valid until dspmm_server_release */
    unsigned flags;     /* DSPMM_REQUEST_* */
} dspmm_request;

/* Server-owned C is freed as soon as the request completes (the result is
 * not kept; for load generators and benchmarks). */
#define DSPMM_REQUEST_DISCARD_RESULT 0x1u

typedef struct dspmm_completion {
    uint64_t id;
    int status;             /* dspmm_status_t */
    int batch_size;
    double t_submit, t_admit, t_done;
    const float *C;         /* device result */
} dspmm_completion;

typedef struct dspmm_server_stats {
    int64_t submitted, completed, rejected, batches;
    int64_t cache_hits, cache_misses, evictions;
    size_t cache_used, cache_peak, operand_peak;
    double mean_batch;
} dspmm_server_stats;

DSPMM_API double dspmm_server_now(void);
DSPMM_API dspmm_status_t dspmm_server_create(dspmm_server_t *s, const dspmm_server_config *cfg);
DSPMM_API dspmm_status_t dspmm_server_destroy(dspmm_server_t s);
/* Registers a CSR matrix (host arrays are copied); it is uploaded when first
 * needed, or now with `preload`. */
DSPMM_API dspmm_status_t dspmm_server_add_matrix(dspmm_server_t s, int nr, int nc, int nnz,
                                                 const int *h_rowPtr, const int *h_colIdx,
                                                 const float *h_values, int preload,
                                                 int *matrix_id);
/* Non-blocking.  DSPMM_STATUS_ADMISSION_REJECTED if the request can never
 * fit (its matrix exceeds cache_bytes or its operands exceed operand_bytes). */
DSPMM_API dspmm_status_t dspmm_server_submit(dspmm_server_t s, const dspmm_request *req);
/* Submits req[i] at time t0 + at[i] (at[] ascending, seconds), from the
 * calling thread, then waits until all of them have completed. */
DSPMM_API dspmm_status_t dspmm_server_replay(dspmm_server_t s, const dspmm_request *req,
                                             const double *at, int n);
/* Completions not yet returned; returns how many were written (<= max). */
DSPMM_API int dspmm_server_poll(dspmm_server_t s, dspmm_completion *out, int max);
DSPMM_API dspmm_status_t dspmm_server_release(dspmm_server_t s, uint64_t id);
DSPMM_API dspmm_status_t dspmm_server_drain(dspmm_server_t s);
DSPMM_API dspmm_status_t dspmm_server_get_stats(dspmm_server_t s, dspmm_server_stats *st);

/* ---- Matrix-Market input (host) ------------------------------------------- */

/* A host COO matrix sorted by (row, column); arrays are malloc'ed by the
 * library and released with dspmm_coo_free. */
typedef struct dspmm_coo {
    int nr, nc, nnz;
    int *rowIdx, *colIdx;
    float *values;
} dspmm_coo;

#define DSPMM_MTX_KEEP_DIAGONAL 0x1u  /* symmetric files: keep diagonal entries */
#define DSPMM_MTX_ALIGN_VALUES  0x2u  /* symmetric files: values follow their entries */

/* Reads a coordinate Matrix-Market file (real, integer or pattern; general or
 * symmetric; 1-based indices).  Replaces readMtx<int,float>
 * (util/spmatrix.hpp:286-333) with mm_read_banner / mm_read_mtx_crd_size
 * (util/mmio.hpp:215-336), called at spmvspmm_benchmark.cu:349.  With
 * flags == 0 the result equals the reference's for the same file: symmetric
 * files are mirrored, duplicates removed, diagonal entries dropped, and the
 * compaction after a dropped entry moves indices but not values
 * (util/spmatrix.hpp:264-283); the two flags above change the last two.
 * Complex files: DSPMM_STATUS_NOT_SUPPORTED; malformed files, dense arrays
 * and out-of-range indices: DSPMM_STATUS_INVALID_VALUE. */
DSPMM_API dspmm_status_t dspmm_mtx_read(const char *path, unsigned flags, dspmm_coo *out);
DSPMM_API void dspmm_coo_free(dspmm_coo *coo);
/* rowPtr[nr + 1] from row indices sorted ascending; replaces
 * generateCsrRowPtrArray (spmvspmm_benchmark.cu:41-76).  Unsorted or
 * out-of-range rows: DSPMM_STATUS_INVALID_VALUE. */
DSPMM_API dspmm_status_t dspmm_csr_rowptr(int nr, int nnz, const int *rowIdx, int *rowPtr);

#ifdef __cplusplus
}
#endif
#endif /* DSPMM_H */
