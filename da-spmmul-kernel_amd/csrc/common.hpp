// common.hpp -- shared device helpers: vector loads/stores along N,
// BF16 widening, block -> XCD remapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wave64.hpp"

namespace dspmm {

constexpr int kBlock = 256;       // 4 wavefronts per workgroup
constexpr int kNumXcd = 8;        // MI355X: 8 XCDs, workgroups dispatched round-robin

// Workgroup b is dispatched to XCD (b % 8).  Renumber so that each XCD
// works on a contiguous range of logical blocks: neighbouring rows of A
// (which share columns of B in banded/stencil matrices) then hit the same
// L2.  A bijection on [0, nblocks) for any nblocks.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    const int per = nblocks / kNumXcd, rem = nblocks % kNumXcd;
    const int x = b % kNumXcd, i = b / kNumXcd;
    return x * per + (x < rem ? x : rem) + i;
}

// ---- dense FP32 vectors along N ----------------------------------------
template <int VW>
__device__ __forceinline__ void load_vec(float (&r)[VW], const float *p) {
    static_assert(VW == 1 || VW == 2 || VW == 4 || VW == 8, "VW");
    if constexpr (VW == 8) {
        const float4 t0 = reinterpret_cast<const float4 *>(p)[0];
        const float4 t1 = reinterpret_cast<const float4 *>(p)[1];
        r[0] = t0.x; r[1] = t0.y; r[2] = t0.z; r[3] = t0.w;
        r[4] = t1.x; r[5] = t1.y; r[6] = t1.z; r[7] = t1.w;
    } else if constexpr (VW == 4) {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        r[0] = t.x; r[1] = t.y; r[2] = t.z; r[3] = t.w;
    } else if constexpr (VW == 2) {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        r[0] = t.x; r[1] = t.y;
    } else {
        r[0] = *p;
    }
}

template <int VW>
__device__ __forceinline__ void store_vec(float *p, const float (&r)[VW]) {
    if constexpr (VW == 8) {
        reinterpret_cast<float4 *>(p)[0] = make_float4(r[0], r[1], r[2], r[3]);
        reinterpret_cast<float4 *>(p)[1] = make_float4(r[4], r[5], r[6], r[7]);
    } else if constexpr (VW == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(r[0], r[1], r[2], r[3]);
    } else if constexpr (VW == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(r[0], r[1]);
    } else {
        *p = r[0];
    }
}

// Streaming store (output is written once and not re-read by the kernel):
// keeps C from evicting the gathered dense input from L2 / MALL.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
template <int VW>
__device__ __forceinline__ void store_vec_nt(float *p, const float (&r)[VW]) {
    if constexpr (VW == 8) {
        nt_f4 t0 = {r[0], r[1], r[2], r[3]};
        nt_f4 t1 = {r[4], r[5], r[6], r[7]};
        __builtin_nontemporal_store(t0, reinterpret_cast<nt_f4 *>(p));
        __builtin_nontemporal_store(t1, reinterpret_cast<nt_f4 *>(p) + 1);
    } else if constexpr (VW == 4) {
        nt_f4 t = {r[0], r[1], r[2], r[3]};
        __builtin_nontemporal_store(t, reinterpret_cast<nt_f4 *>(p));
    } else if constexpr (VW == 2) {
        nt_f2 t = {r[0], r[1]};
        __builtin_nontemporal_store(t, reinterpret_cast<nt_f2 *>(p));
    } else {
        __builtin_nontemporal_store(r[0], p);
    }
}

template <int VW>
__device__ __forceinline__ void atomic_add_vec(float *p, const float (&r)[VW]) {
#pragma unroll
    for (int k = 0; k < VW; k++) unsafeAtomicAdd(p + k, r[k]);
}

// ---- dense BF16 vectors along N (FP32 widening is a 16-bit shift) -------
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

template <int VW>
__device__ __forceinline__ void load_vec_bf16(float (&r)[VW], const uint16_t *p) {
    static_assert(VW == 1 || VW == 2 || VW == 4 || VW == 8, "VW");
    if constexpr (VW == 8) {
        const uint4 t = *reinterpret_cast<const uint4 *>(p);
        r[0] = bf16_lo(t.x); r[1] = bf16_hi(t.x); r[2] = bf16_lo(t.y); r[3] = bf16_hi(t.y);
        r[4] = bf16_lo(t.z); r[5] = bf16_hi(t.z); r[6] = bf16_lo(t.w); r[7] = bf16_hi(t.w);
    } else if constexpr (VW == 4) {
        const uint2 t = *reinterpret_cast<const uint2 *>(p);
        r[0] = bf16_lo(t.x); r[1] = bf16_hi(t.x); r[2] = bf16_lo(t.y); r[3] = bf16_hi(t.y);
    } else if constexpr (VW == 2) {
        const uint32_t t = *reinterpret_cast<const uint32_t *>(p);
        r[0] = bf16_lo(t); r[1] = bf16_hi(t);
    } else {
        r[0] = __uint_as_float((uint32_t)(*p) << 16);
    }
}

// Typed dense-input accessor: B element type is float or uint16_t (BF16).
template <typename TB, int VW>
__device__ __forceinline__ void load_b(float (&r)[VW], const TB *p) {
    if constexpr (sizeof(TB) == 4) load_vec<VW>(r, reinterpret_cast<const float *>(p));
    else load_vec_bf16<VW>(r, reinterpret_cast<const uint16_t *>(p));
}

// Streaming (read-once) scalar loads for the sparse operand.
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }

}  // namespace dspmm
