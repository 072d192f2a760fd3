// api_compat.cpp -- C-ABI: the 12 reference-named launchers, the
// cudaSpmm-shaped dispatcher, device memory / streams / timers (CUArray and
// GpuTimer replacements), and the benchmark / dry-run loops of
// SpmvSpmmProblem.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace dspmm {

namespace {
thread_local int g_last_hip_error = 0;
}

dspmm_status_t hip_status(hipError_t e) {
    if (e == hipSuccess) return DSPMM_STATUS_SUCCESS;
    g_last_hip_error = (int)e;
    if (e == hipErrorOutOfMemory) return DSPMM_STATUS_ALLOC_FAILED;
    return DSPMM_STATUS_HIP_ERROR;
}

namespace {

// Argument checks shared by the launchers.  Pointers may be NULL only
// where the corresponding size makes them unused.
dspmm_status_t check_common(int nr, int nc, int nnz, int nv, const int *rowPtr,
                            const int *rowIdx, const int *colIdx, const float *values,
                            const float *B, const float *C, bool coo) {
    if (nr < 0 || nc < 0 || nnz < 0 || nv < 1) return DSPMM_STATUS_INVALID_VALUE;
    if (nr > 0 && !C) return DSPMM_STATUS_INVALID_VALUE;
    if (!coo && nr > 0 && !rowPtr) return DSPMM_STATUS_INVALID_VALUE;
    if (nnz > 0 && (!colIdx || !values || !B)) return DSPMM_STATUS_INVALID_VALUE;
    if (coo && nnz > 0 && !rowIdx) return DSPMM_STATUS_INVALID_VALUE;
    return DSPMM_STATUS_SUCCESS;
}

bool is_coo(int which) {
    return which == K_SPMV_COO_SCALAR || which == K_SPMV_COO_VECTOR ||
           which == K_SPMM_COO_SCALAR_ROW || which == K_SPMM_COO_VECTOR_ROW ||
           which == K_SPMM_COO_SCALAR_COL || which == K_SPMM_COO_VECTOR_COL;
}

bool is_col(int which) {
    return which == K_SPMM_CSR_SCALAR_COL || which == K_SPMM_CSR_VECTOR_COL ||
           which == K_SPMM_COO_SCALAR_COL || which == K_SPMM_COO_VECTOR_COL;
}

// One reference-named launcher: packed operands (ld = nv row-major; ld =
// nc / nr column-major; ld = 1 for the SpMV kernels, whose nv==1 case is
// the same in both layouts).
dspmm_status_t run_named(int which, int nr, int nc, int nnz, int nv, const int *rowPtr,
                         const int *rowIdx, const int *colIdx, const float *values,
                         const float *B, float *C, dspmm_stream_t stream) {
    dspmm_status_t st = check_common(nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, B, C,
                                     is_coo(which));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    switch (which) {
    case K_SPMV_CSR_SCALAR:
    case K_SPMV_COO_SCALAR:
        if (nv != 1) return DSPMM_STATUS_NOT_SUPPORTED;
        break;
    case K_SPMV_CSR_VECTOR:
    case K_SPMV_COO_VECTOR:
        if (nv != 1 && nv != 2 && nv != 4) return DSPMM_STATUS_NOT_SUPPORTED;
        break;
    default:
        break;
    }
    int64_t ldb, ldc;
    if (is_col(which)) { ldb = nc; ldc = nr; }
    else { ldb = nv; ldc = nv; }
    return launch_compat(which, nr, nc, nnz, nv, ldb, ldc, rowPtr, rowIdx, colIdx, values, B, C,
                         as_hip(stream), 0);
}

// spmvspmm.h:72-197: (alg, layout, nv) -> reference kernel.
int select_kernel(SPMV_SPMM_ALG alg, DenseLayout layout, int nv) {
    if (layout == DENSE_ROW_MAJOR) {
        switch (alg) {
        case ALG_CSR_SCALAR: return nv == 1 ? K_SPMV_CSR_SCALAR : K_SPMM_CSR_SCALAR_ROW;
        case ALG_COO_SCALAR: return nv == 1 ? K_SPMV_COO_SCALAR : K_SPMM_COO_SCALAR_ROW;
        case ALG_CSR_VECTOR: return nv <= 4 ? K_SPMV_CSR_VECTOR : K_SPMM_CSR_VECTOR_ROW;
        case ALG_COO_VECTOR: return nv <= 4 ? K_SPMV_COO_VECTOR : K_SPMM_COO_VECTOR_ROW;
        }
    } else if (layout == DENSE_COL_MAJOR) {
        switch (alg) {
        case ALG_CSR_SCALAR: return nv == 1 ? K_SPMV_CSR_SCALAR : K_SPMM_CSR_SCALAR_COL;
        case ALG_COO_SCALAR: return nv == 1 ? K_SPMV_COO_SCALAR : K_SPMM_COO_SCALAR_COL;
        case ALG_CSR_VECTOR: return nv == 1 ? K_SPMV_CSR_VECTOR : K_SPMM_CSR_VECTOR_COL;
        case ALG_COO_VECTOR: return nv == 1 ? K_SPMV_COO_VECTOR : K_SPMM_COO_VECTOR_COL;
        }
    }
    return -1;
}

}  // namespace
}  // namespace dspmm

using namespace dspmm;

extern "C" {

const char *dspmm_get_error_string(dspmm_status_t s) {
    switch (s) {
    case DSPMM_STATUS_SUCCESS: return "success";
    case DSPMM_STATUS_INVALID_VALUE: return "invalid value";
    case DSPMM_STATUS_NOT_SUPPORTED: return "not supported";
    case DSPMM_STATUS_ALLOC_FAILED: return "allocation failed";
    case DSPMM_STATUS_HIP_ERROR: return "HIP runtime error";
    case DSPMM_STATUS_WORKSPACE_TOO_SMALL: return "workspace too small";
    case DSPMM_STATUS_ADMISSION_REJECTED: return "admission rejected (HBM budget)";
    case DSPMM_STATUS_NOT_PREPARED:
        return "not prepared (call dspmm_matrix_prepare with these arguments first)";
    }
    return "unknown status";
}

int dspmm_last_hip_error(void) { return g_last_hip_error; }

const char *dspmm_version(void) { return "dspmm 0.1 (gfx950)"; }

#define DSPMM_NAMED(fn, K)                                                                   \
    dspmm_status_t fn(int nr, int nc, int nnz, int nv, const int *rowPtr, const int *rowIdx, \
                      const int *colIdx, const float *values, const float *dnInput,          \
                      float *dnOutput, dspmm_stream_t stream) {                              \
        return run_named(K, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput,       \
                         dnOutput, stream);                                                  \
    }

DSPMM_NAMED(spmv_csr_scalar, K_SPMV_CSR_SCALAR)
DSPMM_NAMED(spmv_coo_scalar, K_SPMV_COO_SCALAR)
DSPMM_NAMED(spmv_csr_vector, K_SPMV_CSR_VECTOR)
DSPMM_NAMED(spmv_coo_vector, K_SPMV_COO_VECTOR)
DSPMM_NAMED(spmm_csr_scalar_row, K_SPMM_CSR_SCALAR_ROW)
DSPMM_NAMED(spmm_csr_vector_row, K_SPMM_CSR_VECTOR_ROW)
DSPMM_NAMED(spmm_coo_scalar_row, K_SPMM_COO_SCALAR_ROW)
DSPMM_NAMED(spmm_coo_vector_row, K_SPMM_COO_VECTOR_ROW)
DSPMM_NAMED(spmm_csr_scalar_col, K_SPMM_CSR_SCALAR_COL)
DSPMM_NAMED(spmm_csr_vector_col, K_SPMM_CSR_VECTOR_COL)
DSPMM_NAMED(spmm_coo_scalar_col, K_SPMM_COO_SCALAR_COL)
DSPMM_NAMED(spmm_coo_vector_col, K_SPMM_COO_VECTOR_COL)
#undef DSPMM_NAMED

dspmm_status_t dspmm_spmm(SPMV_SPMM_ALG alg, DenseLayout layout, int nr, int nc, int nnz, int nv,
                          const int *rowPtr, const int *rowIdx, const int *colIdx,
                          const float *values, const float *dnInput, float *dnOutput,
                          dspmm_stream_t stream) {
    if (nv < 1) return DSPMM_STATUS_INVALID_VALUE;
    if ((nv & (nv - 1)) != 0) return DSPMM_STATUS_NOT_SUPPORTED;   // spmvspmm.h:209
    const int k = select_kernel(alg, layout, nv);
    if (k < 0) return DSPMM_STATUS_INVALID_VALUE;
    return run_named(k, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput, dnOutput,
                     stream);
}

// ---- memory, streams, device --------------------------------------------

dspmm_status_t dspmm_device_alloc(void **ptr, size_t bytes) {
    if (!ptr) return DSPMM_STATUS_INVALID_VALUE;
    *ptr = nullptr;
    if (bytes == 0) return DSPMM_STATUS_SUCCESS;
    return hip_status(hipMalloc(ptr, bytes));
}

dspmm_status_t dspmm_device_free(void *ptr) {
    if (!ptr) return DSPMM_STATUS_SUCCESS;
    return hip_status(hipFree(ptr));
}

dspmm_status_t dspmm_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return DSPMM_STATUS_SUCCESS;
    if (!dst || !src) return DSPMM_STATUS_INVALID_VALUE;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

dspmm_status_t dspmm_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return DSPMM_STATUS_SUCCESS;
    if (!dst || !src) return DSPMM_STATUS_INVALID_VALUE;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

dspmm_status_t dspmm_memcpy_d2d(void *dst, const void *src, size_t bytes, dspmm_stream_t stream) {
    if (bytes == 0) return DSPMM_STATUS_SUCCESS;
    if (!dst || !src) return DSPMM_STATUS_INVALID_VALUE;
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_hip(stream)));
}

dspmm_status_t dspmm_stream_create(dspmm_stream_t *stream) {
    if (!stream) return DSPMM_STATUS_INVALID_VALUE;
    hipStream_t s = nullptr;
    dspmm_status_t st = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<dspmm_stream_t>(s);
    return st;
}

dspmm_status_t dspmm_stream_destroy(dspmm_stream_t stream) {
    if (!stream) return DSPMM_STATUS_SUCCESS;
    return hip_status(hipStreamDestroy(as_hip(stream)));
}

dspmm_status_t dspmm_stream_synchronize(dspmm_stream_t stream) {
    return hip_status(hipStreamSynchronize(as_hip(stream)));
}

dspmm_status_t dspmm_device_synchronize(void) { return hip_status(hipDeviceSynchronize()); }

dspmm_status_t dspmm_get_device_info(dspmm_device_info *info) {
    if (!info) return DSPMM_STATUS_INVALID_VALUE;
    std::memset(info, 0, sizeof(*info));
    int dev = 0;
    dspmm_status_t st = hip_status(hipGetDevice(&dev));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    hipDeviceProp_t p;
    st = hip_status(hipGetDeviceProperties(&p, dev));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    std::strncpy(info->name, p.name, sizeof(info->name) - 1);
    std::strncpy(info->arch, p.gcnArchName, sizeof(info->arch) - 1);
    info->compute_units = p.multiProcessorCount;
    info->wave_size = p.warpSize;
    info->clock_khz = p.clockRate;
    info->l2_bytes = p.l2CacheSize;
    info->lds_per_block = p.sharedMemPerBlock;
    size_t fr = 0, tot = 0;
    st = hip_status(hipMemGetInfo(&fr, &tot));
    info->total_mem = tot;
    info->free_mem = fr;
    return st;
}

// ---- event timer ----------------------------------------------------------

struct dspmm_timer_s {
    std::vector<hipEvent_t> ev;
    int used = 0;
};

dspmm_status_t dspmm_timer_create(dspmm_timer_t *t, int max_marks) {
    if (!t || max_marks < 2) return DSPMM_STATUS_INVALID_VALUE;
    auto *tm = new dspmm_timer_s;
    tm->ev.resize((size_t)max_marks, nullptr);
    for (auto &e : tm->ev) {
        dspmm_status_t st = hip_status(hipEventCreate(&e));
        if (st != DSPMM_STATUS_SUCCESS) {
            for (auto &x : tm->ev) if (x) hipEventDestroy(x);
            delete tm;
            return st;
        }
    }
    *t = tm;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_timer_destroy(dspmm_timer_t t) {
    if (!t) return DSPMM_STATUS_SUCCESS;
    for (auto &e : t->ev) hipEventDestroy(e);
    delete t;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_timer_reset(dspmm_timer_t t) {
    if (!t) return DSPMM_STATUS_INVALID_VALUE;
    t->used = 0;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_timer_mark(dspmm_timer_t t, dspmm_stream_t stream) {
    if (!t) return DSPMM_STATUS_INVALID_VALUE;
    if (t->used >= (int)t->ev.size()) return DSPMM_STATUS_INVALID_VALUE;
    return hip_status(hipEventRecord(t->ev[(size_t)t->used++], as_hip(stream)));
}

dspmm_status_t dspmm_timer_intervals(dspmm_timer_t t, float *ms, int n) {
    if (!t || t->used < 2 || n < 0 || n > t->used - 1) return DSPMM_STATUS_INVALID_VALUE;
    dspmm_status_t st = hip_status(hipEventSynchronize(t->ev[(size_t)t->used - 1]));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    for (int i = 0; i < n; i++) {
        st = hip_status(hipEventElapsedTime(&ms[i], t->ev[(size_t)i], t->ev[(size_t)i + 1]));
        if (st != DSPMM_STATUS_SUCCESS) return st;
    }
    return DSPMM_STATUS_SUCCESS;
}

// ---- SpmvSpmmProblem::benchmark / dry_run -----------------------------------

dspmm_status_t dspmm_benchmark_compat(SPMV_SPMM_ALG alg, DenseLayout layout, int nr, int nc,
                                      int nnz, int nv, const int *rowPtr, const int *rowIdx,
                                      const int *colIdx, const float *values,
                                      const float *dnInput, float *dnOutput, int warmup,
                                      int reps, float *ms_mean, float *ms_per_launch,
                                      dspmm_stream_t stream) {
    if (reps < 1 || warmup < 0) return DSPMM_STATUS_INVALID_VALUE;
    dspmm_status_t st = DSPMM_STATUS_SUCCESS;
    for (int i = 0; i < warmup && st == DSPMM_STATUS_SUCCESS; i++)
        st = dspmm_spmm(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput,
                        dnOutput, stream);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    st = hip_status(hipStreamSynchronize(as_hip(stream)));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    dspmm_timer_t t = nullptr;
    st = dspmm_timer_create(&t, reps + 1);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    st = dspmm_timer_mark(t, stream);
    for (int i = 0; i < reps && st == DSPMM_STATUS_SUCCESS; i++) {
        st = dspmm_spmm(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput,
                        dnOutput, stream);
        if (st == DSPMM_STATUS_SUCCESS) st = dspmm_timer_mark(t, stream);
    }
    if (st == DSPMM_STATUS_SUCCESS) {
        std::vector<float> ms((size_t)reps);
        st = dspmm_timer_intervals(t, ms.data(), reps);
        if (st == DSPMM_STATUS_SUCCESS) {
            double sum = 0;
            for (float x : ms) sum += x;
            if (ms_mean) *ms_mean = (float)(sum / reps);
            if (ms_per_launch) std::copy(ms.begin(), ms.end(), ms_per_launch);
        }
    }
    dspmm_timer_destroy(t);
    return st;
}

dspmm_status_t dspmm_dry_run_compat(SPMV_SPMM_ALG alg, DenseLayout layout, int nr, int nc,
                                    int nnz, int nv, const int *rowPtr, const int *rowIdx,
                                    const int *colIdx, const float *values,
                                    const float *dnInput, float *dnOutput, int warmup,
                                    dspmm_stream_t stream) {
    if (warmup < 0) return DSPMM_STATUS_INVALID_VALUE;
    dspmm_status_t st = DSPMM_STATUS_SUCCESS;
    for (int i = 0; i < warmup && st == DSPMM_STATUS_SUCCESS; i++)
        st = dspmm_spmm(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput,
                        dnOutput, stream);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    st = hip_status(hipStreamSynchronize(as_hip(stream)));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    roctxProfilerResume(0);
    roctxRangePushA("dspmm_dry_run");
    st = dspmm_spmm(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values, dnInput,
                    dnOutput, stream);
    dspmm_status_t st2 = hip_status(hipStreamSynchronize(as_hip(stream)));
    roctxRangePop();
    roctxProfilerPause(0);
    return st != DSPMM_STATUS_SUCCESS ? st : st2;
}

}  // extern "C"
