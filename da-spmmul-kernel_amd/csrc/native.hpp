// native.hpp -- host-side interface of the native kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace dspmm {

// Arguments of one nnz-tiled SpMM launch (plus its split-row fix-up).
struct NbLaunch {
    const int4 *tiles = nullptr;   // {row_begin, row_end | -1-slot, p_begin, p_end}
    int ntiles = 0;
    const int4 *splits = nullptr;  // {row, first_slot, nchunks, 0}
    int nsplits = 0;
    const int *rowPtr = nullptr;
    const int *colIdx = nullptr;
    const float *values = nullptr;
    const void *B = nullptr;
    int64_t ldb = 0;
    bool bf16 = false;
    float *C = nullptr;
    int64_t ldc = 0;
    float *carry = nullptr;        // nslots x ldcarry floats
    int64_t ldcarry = 0;
    int nv = 0;
    int vw = 1;
    int lanes = 1;
    int slab_w = 1;
    int nslabs = 1;
    int S = 64;
    int unroll = 8;
    hipEvent_t mid_event = nullptr;  // recorded between the main kernel and the fix-up
};

dspmm_status_t launch_nnz_tiled(const NbLaunch &a, hipStream_t st);
size_t nnz_tiled_lds_bytes(int lanes, int S);
dspmm_status_t launch_fill_zero(void *p, size_t bytes, hipStream_t st);
dspmm_status_t launch_f32_to_bf16(const float *src, uint16_t *dst, int64_t n, hipStream_t st);

}  // namespace dspmm
