// internal.hpp -- host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dspmm.h"

namespace dspmm {

// Records `e` as this thread's last HIP error and maps it to a status.
dspmm_status_t hip_status(hipError_t e);
// Status of the most recent launch on this thread.
inline dspmm_status_t launch_status() { return hip_status(hipGetLastError()); }

inline hipStream_t as_hip(dspmm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
inline int pow2_ceil(int x) { int p = 1; while (p < x) p <<= 1; return p; }
inline int pow2_floor(int x) { int p = 1; while (p * 2 <= x) p <<= 1; return p; }
inline bool aligned(const void *p, int bytes) { return ((uintptr_t)p % (uintptr_t)bytes) == 0; }

// Vector width along N for row-major operands: the widest of 4/2/1 floats
// (8 also for BF16) that divides nv and both leading dimensions and keeps
// every vector access naturally aligned.
inline int choose_vw(int nv, int64_t ldb, int64_t ldc, const void *B, const void *C,
                     int elem_b_bytes, int max_vw) {
    for (int vw = max_vw; vw >= 2; vw /= 2) {
        if (nv % vw || ldb % vw || ldc % vw) continue;
        if (!aligned(B, vw * elem_b_bytes) || !aligned(C, vw * 4)) continue;
        return vw;
    }
    return 1;
}

// Lane-group shape covering nv columns with `vw` floats per lane: lanes is a
// power of two <= 64, slab = lanes*vw columns per grid.y slice.
struct Slab {
    int vw, lanes, width, count;
};
inline Slab make_slab(int nv, int vw) {
    Slab s;
    s.vw = vw;
    s.lanes = pow2_ceil(div_up(nv, vw));
    if (s.lanes > 64) s.lanes = 64;
    s.width = s.lanes * vw;
    s.count = div_up(nv, s.width);
    return s;
}

// Compat launchers shared with the dispatcher (defined in compat_kernels.hip).
dspmm_status_t launch_compat(int which, int nr, int nc, int nnz, int nv, int64_t ldb,
                             int64_t ldc, const int *rowPtr, const int *rowIdx,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t stream, int group_hint);

enum CompatKernel {
    K_SPMV_CSR_SCALAR = 0,
    K_SPMV_COO_SCALAR,
    K_SPMV_CSR_VECTOR,
    K_SPMV_COO_VECTOR,
    K_SPMM_CSR_SCALAR_ROW,
    K_SPMM_CSR_VECTOR_ROW,
    K_SPMM_COO_SCALAR_ROW,
    K_SPMM_COO_VECTOR_ROW,
    K_SPMM_CSR_SCALAR_COL,
    K_SPMM_CSR_VECTOR_COL,
    K_SPMM_COO_SCALAR_COL,
    K_SPMM_COO_VECTOR_COL,
    K_COUNT
};

}  // namespace dspmm
