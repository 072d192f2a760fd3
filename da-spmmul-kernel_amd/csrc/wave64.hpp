// wave64.hpp -- cross-lane primitives for 64-wide CDNA4 wavefronts.
//
// The reference reduces with 32-lane warp shuffles (helper.h:269-281:
// SHFL_DOWN_REDUCE, a 5-step tree, and SEG_SHFL_SCAN, a 5-step segmented
// suffix scan).  Here a wavefront has 64 lanes, and lanes are exchanged
// with DPP (within 16-lane rows), ds_swizzle (within 32-lane halves) and
// v_permlane32_swap (across halves, gfx950).
//
// Lane groups: a "group" is L consecutive lanes that together hold one
// slice of a dense row (L is a power of two).  Lanes that hold the same
// column slice in different groups are L apart.  All functions must be
// called with every lane of the wavefront active.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dspmm {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// DPP controls (GFX9 encoding).
constexpr int kDppQuadXor1 = 0xB1;   // quad_perm:[1,0,3,2]  -> lane ^ 1
constexpr int kDppQuadXor2 = 0x4E;   // quad_perm:[2,3,0,1]  -> lane ^ 2
constexpr int kDppRowRor8 = 0x128;   // row_ror:8            -> lane ^ 8 within a row
constexpr int dpp_row_shr(int n) { return 0x110 + n; }   // row_shr:n (1..15)

template <int CTRL>
__device__ __forceinline__ int dpp_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v),
                                                      CTRL, 0xF, 0xF, false));
}

// ds_swizzle in bitmask mode: within each 32-lane half, lane i reads lane
// ((i & AND) | OR) ^ XOR.
constexpr int swz(int and_mask, int or_mask, int xor_mask) {
    return (and_mask & 0x1F) | ((or_mask & 0x1F) << 5) | ((xor_mask & 0x1F) << 10);
}
template <int OFF>
__device__ __forceinline__ int swizzle_i(int v) {
    return __builtin_amdgcn_ds_swizzle(v, OFF);
}
template <int OFF>
__device__ __forceinline__ float swizzle_f(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), OFF));
}

// Value held by lane (i ^ 32): one v_permlane32_swap gives both halves.
__device__ __forceinline__ int other_half_i(int v) {
    auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return lane_id() < 32 ? (int)r[1] : (int)r[0];
}
__device__ __forceinline__ float other_half_f(float v) {
    return __int_as_float(other_half_i(__float_as_int(v)));
}
// Lanes 32..63 get the value of lane (i - 32); lanes 0..31 keep their own.
__device__ __forceinline__ int from_lower_half_i(int v) {
    auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (int)r[0];
}

// Butterfly partner: value held by lane (i ^ S), S a power of two <= 32.
template <int S>
__device__ __forceinline__ float xor_lane(float v) {
    static_assert(S >= 1 && S <= 32 && (S & (S - 1)) == 0, "S must be a power of two <= 32");
    if constexpr (S == 1) return dpp_f<kDppQuadXor1>(v, v);
    else if constexpr (S == 2) return dpp_f<kDppQuadXor2>(v, v);
    else if constexpr (S == 4) return swizzle_f<swz(0x1F, 0, 4)>(v);
    else if constexpr (S == 8) return dpp_f<kDppRowRor8>(v, v);
    else if constexpr (S == 16) return swizzle_f<swz(0x1F, 0, 16)>(v);
    else return other_half_f(v);
}

// Sum over the lanes {i ^ (k*FROM) : k} inside an aligned block of TO lanes,
// i.e. butterfly strides FROM, 2*FROM, ..., TO/2.  Every lane of the block
// ends with the block's sum for its residue class mod FROM.
template <int FROM, int TO>
__device__ __forceinline__ float butterfly_sum(float v) {
    if constexpr (FROM < TO) {
        v += xor_lane<FROM>(v);
        return butterfly_sum<FROM * 2, TO>(v);
    } else {
        return v;
    }
}

// Sum over an aligned group of L consecutive lanes (all lanes get it).
template <int L>
__device__ __forceinline__ float group_sum(float v) { return butterfly_sum<1, L>(v); }

// Segmented inclusive scan over the G = 64/L groups of a wavefront, in
// group order.  `key` is the segment id of the group (equal in all its
// lanes) and must be non-decreasing with the group index; `v` are the
// NV values each lane holds for its column slice.  Afterwards the last
// group of every segment holds the segment's sum.
//
// Two levels: Hillis-Steele steps with DPP row_shr inside 16-lane rows,
// then row 1 (and row 3) take the last group of row 0 (row 2) through
// ds_swizzle, then rows 2-3 take the last group of row 1 through
// v_permlane32_swap + ds_swizzle.  Because keys are sorted, equal keys at
// two lanes imply one segment between them, so a key compare is the flag.
template <int L, int NV>
__device__ __forceinline__ void segmented_scan(float (&v)[NV], int key) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L must be a power of two");
    if constexpr (L < 64) {
        const int ln = lane_id();
        // Level 1: within 16-lane rows.
        if constexpr (L <= 1) {
            const int k1 = dpp_i<dpp_row_shr(1)>(INT32_MIN, key);
#pragma unroll
            for (int j = 0; j < NV; j++) { float t = dpp_f<dpp_row_shr(1)>(0.f, v[j]); if (k1 == key) v[j] += t; }
        }
        if constexpr (L <= 2) {
            const int k2 = dpp_i<dpp_row_shr(2)>(INT32_MIN, key);
#pragma unroll
            for (int j = 0; j < NV; j++) { float t = dpp_f<dpp_row_shr(2)>(0.f, v[j]); if (k2 == key) v[j] += t; }
        }
        if constexpr (L <= 4) {
            const int k4 = dpp_i<dpp_row_shr(4)>(INT32_MIN, key);
#pragma unroll
            for (int j = 0; j < NV; j++) { float t = dpp_f<dpp_row_shr(4)>(0.f, v[j]); if (k4 == key) v[j] += t; }
        }
        if constexpr (L <= 8) {
            const int k8 = dpp_i<dpp_row_shr(8)>(INT32_MIN, key);
#pragma unroll
            for (int j = 0; j < NV; j++) { float t = dpp_f<dpp_row_shr(8)>(0.f, v[j]); if (k8 == key) v[j] += t; }
        }
        // Level 2: rows 1 and 3 add the last group of rows 0 and 2.
        if constexpr (L <= 16) {
            constexpr int OFF = swz(L - 1, 16 - L, 0);
            const int kk = swizzle_i<OFF>(key);
            const bool take = (ln & 16) && kk == key;
#pragma unroll
            for (int j = 0; j < NV; j++) { float t = swizzle_f<OFF>(v[j]); if (take) v[j] += t; }
        }
        // Level 3: rows 2 and 3 add the last group of row 1.
        {
            constexpr int OFF = swz(L - 1, 32 - L, 0);
            const int kk = swizzle_i<OFF>(from_lower_half_i(key));
            const bool take = (ln & 32) && kk == key;
#pragma unroll
            for (int j = 0; j < NV; j++) {
                float t = __int_as_float(swizzle_i<OFF>(from_lower_half_i(__float_as_int(v[j]))));
                if (take) v[j] += t;
            }
        }
    }
}

// Read lane `src` of a wave-uniform-independent value (result is uniform).
__device__ __forceinline__ int read_lane(int v, int src) {
    return __builtin_amdgcn_readlane(v, src);
}

}  // namespace dspmm
