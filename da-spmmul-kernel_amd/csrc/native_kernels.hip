// native_kernels.hip -- the native load-balanced SpMM path and utility
// kernels.
//
// Nnz-tiled SpMM (row-major, overwrite, deterministic).  A host-built plan
// cuts the matrix into tiles of at most S nonzeros that start and end on
// row boundaries; a row longer than S is cut into S-sized chunks whose
// partial sums go to a carry buffer and are combined, in a fixed order, by
// a second small kernel.  Every tile is owned by a group of L lanes x VW
// floats that sums each of its rows sequentially, in nonzero order, with a
// separately rounded multiply and add -- so every row that is not split
// equals the reference host loop (spmvspmm_benchmark.cu:11-39) bit for
// bit, with FP32 or BF16-stored B.  A wavefront owns G = 64/L consecutive
// tiles; it stages their (col, val) and row offsets through LDS with
// coalesced loads, then issues UNR B-row gathers per group before summing
// them.  Stores of C are non-temporal so C does not evict B from L2/MALL.
#include <algorithm>
#include <climits>

#include "common.hpp"
#include "internal.hpp"
#include "native.hpp"

namespace dspmm {

// LDS position of staged element i: one pad word every 16 elements, so that
// groups reading their tiles' k-th element at 16-element strides spread
// over banks.
__device__ __forceinline__ int skew(int i) { return i + (i >> 4); }

template <typename TB, int VW, int L, int UNR>
__global__ __launch_bounds__(kBlock) void k_nnz_tiled(
    const int4 *__restrict__ tiles, int ntiles, const int *__restrict__ rowPtr,
    const int *__restrict__ colIdx, const float *__restrict__ values, const TB *__restrict__ B,
    int64_t ldb, float *__restrict__ C, int64_t ldc, float *__restrict__ carry, int64_t ldcarry,
    int nv, int slab_w, int S) {
    constexpr int G = kWave / L;
    extern __shared__ int smem[];
    const int wib = (int)threadIdx.x / kWave;
    const int ln = lane_id();
    const int g = ln / L;
    const int col0 = blockIdx.y * slab_w + (ln % L) * VW;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int t0 = (b * (kBlock / kWave) + wib) * G;

    // Per-wave LDS: cols and vals (skewed), then row ends.
    const int cap = G * S;
    const int capk = skew(cap);
    int *s_col = smem + wib * (2 * capk + cap + 1);
    float *s_val = reinterpret_cast<float *>(s_col + capk);
    int *s_rp = s_col + 2 * capk;

    int P0 = 0, R0 = 0;
    if (t0 < ntiles) {
        const int tl_i = min(t0 + G, ntiles) - 1;
        const int4 tf = tiles[t0];
        const int4 tl = tiles[tl_i];
        P0 = tf.z;
        R0 = tf.x;
        const int P1 = tl.w;
        const int R1 = tl.y >= 0 ? tl.y : tl.x + 1;
        for (int i = ln; i < P1 - P0; i += kWave) {
            s_col[skew(i)] = ld_stream(colIdx + P0 + i);
            s_val[skew(i)] = ld_stream(values + P0 + i);
        }
        for (int i = ln; i < R1 - R0; i += kWave) s_rp[i] = rowPtr[R0 + 1 + i];
    }
    __syncthreads();

    const int t = t0 + g;
    if (t >= ntiles || col0 >= nv) return;
    const int4 tile = tiles[t];
    const int pb = tile.z, pe = tile.w;
    float acc[VW];
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = 0.f;

    if (tile.y < 0) {
        // Chunk of a split row: sum it and leave it in its carry slot.
        for (int p = pb; p < pe; p += UNR) {
            float bv[UNR][VW];
#pragma unroll
            for (int u = 0; u < UNR; u++)
                if (p + u < pe) load_b<TB, VW>(bv[u], B + (int64_t)s_col[skew(p + u - P0)] * ldb + col0);
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                if (p + u < pe) {
                    const float a = s_val[skew(p + u - P0)];
#pragma unroll
                    for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * bv[u][k];
                }
            }
        }
        const int slot = -1 - tile.y;
        store_vec<VW>(carry + (int64_t)slot * ldcarry + col0, acc);
        return;
    }

    int r = tile.x;
    const int re = tile.y;
    int next_end = s_rp[r - R0];                  // rowPtr[r + 1]
    for (int p = pb; p < pe; p += UNR) {
        float bv[UNR][VW];
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (p + u < pe) load_b<TB, VW>(bv[u], B + (int64_t)s_col[skew(p + u - P0)] * ldb + col0);
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int q = p + u;
            if (q < pe) {
                while (q >= next_end) {           // row r is complete (or empty)
                    store_vec_nt<VW>(C + (int64_t)r * ldc + col0, acc);
#pragma unroll
                    for (int k = 0; k < VW; k++) acc[k] = 0.f;
                    ++r;
                    next_end = s_rp[r - R0];
                }
                const float a = s_val[skew(q - P0)];
#pragma unroll
                for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * bv[u][k];
            }
        }
    }
    for (; r < re; ++r) {                          // last row, then trailing empty rows
        store_vec_nt<VW>(C + (int64_t)r * ldc + col0, acc);
#pragma unroll
        for (int k = 0; k < VW; k++) acc[k] = 0.f;
    }
}

// Combines the chunks of each split row: one wavefront per row; group g
// sums chunks g, g+G, ... in order, then a fixed butterfly adds the G
// group sums.  Deterministic for a given plan and shape.
template <int VW, int L>
__global__ __launch_bounds__(kBlock) void k_split_fixup(const int4 *__restrict__ splits,
                                                         int nsplits,
                                                         const float *__restrict__ carry,
                                                         int64_t ldcarry, float *__restrict__ C,
                                                         int64_t ldc, int nv, int slab_w) {
    constexpr int G = kWave / L;
    const int w = (int)(blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (w >= nsplits) return;                        // wave-uniform
    const int ln = lane_id();
    const int g = ln / L;
    const int col0 = blockIdx.y * slab_w + (ln % L) * VW;
    const bool colok = col0 < nv;
    const int4 sp = splits[w];                       // {row, first slot, nchunks, 0}
    float acc[VW];
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = 0.f;
    if (colok) {
        for (int c = g; c < sp.z; c += G) {
            float v[VW];
            load_vec<VW>(v, carry + (int64_t)(sp.y + c) * ldcarry + col0);
#pragma unroll
            for (int k = 0; k < VW; k++) acc[k] = acc[k] + v[k];
        }
    }
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = butterfly_sum<L, kWave>(acc[k]);
    if (g == 0 && colok) store_vec<VW>(C + (int64_t)sp.x * ldc + col0, acc);
}

__global__ void k_fill_zero_u32(uint32_t *__restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}
__global__ void k_fill_zero_u8(uint8_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// FP32 -> BF16, round to nearest even on the bit pattern; NaNs stay NaN.
__global__ void k_f32_to_bf16(const float *__restrict__ src, uint16_t *__restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t u = __float_as_uint(src[i]);
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
            u = (u | 0x00400000u);
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);
        }
        dst[i] = (uint16_t)(u >> 16);
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
size_t nnz_tiled_lds_bytes(int lanes, int S) {
    const int G = kWave / lanes;
    const int cap = G * S;
    const int capk = cap + (cap >> 4);
    return (size_t)(kBlock / kWave) * (size_t)(2 * capk + cap + 1) * sizeof(int);
}

namespace {

template <typename TB, int VW, int L>
void launch_nb(const NbLaunch &a, hipStream_t st) {
    constexpr int G = kWave / L;
    const int waves = div_up(a.ntiles, G);
    dim3 grid(div_up(waves, kBlock / kWave), a.nslabs);
    const size_t lds = nnz_tiled_lds_bytes(L, a.S);
    if (a.unroll <= 4) {
        hipLaunchKernelGGL((k_nnz_tiled<TB, VW, L, 4>), grid, dim3(kBlock), lds, st, a.tiles,
                           a.ntiles, a.rowPtr, a.colIdx, a.values,
                           reinterpret_cast<const TB *>(a.B), a.ldb, a.C, a.ldc, a.carry,
                           a.ldcarry, a.nv, a.slab_w, a.S);
    } else {
        hipLaunchKernelGGL((k_nnz_tiled<TB, VW, L, 8>), grid, dim3(kBlock), lds, st, a.tiles,
                           a.ntiles, a.rowPtr, a.colIdx, a.values,
                           reinterpret_cast<const TB *>(a.B), a.ldb, a.C, a.ldc, a.carry,
                           a.ldcarry, a.nv, a.slab_w, a.S);
    }
    if (a.mid_event) (void)hipEventRecord(a.mid_event, st);
    if (a.nsplits > 0) {
        dim3 fgrid(div_up(a.nsplits, kBlock / kWave), a.nslabs);
        hipLaunchKernelGGL((k_split_fixup<VW, L>), fgrid, dim3(kBlock), 0, st, a.splits,
                           a.nsplits, a.carry, a.ldcarry, a.C, a.ldc, a.nv, a.slab_w);
    }
}

template <typename TB, int VW>
void launch_nb_l(const NbLaunch &a, hipStream_t st) {
    switch (a.lanes) {
    case 1: launch_nb<TB, VW, 1>(a, st); break;
    case 2: launch_nb<TB, VW, 2>(a, st); break;
    case 4: launch_nb<TB, VW, 4>(a, st); break;
    case 8: launch_nb<TB, VW, 8>(a, st); break;
    case 16: launch_nb<TB, VW, 16>(a, st); break;
    case 32: launch_nb<TB, VW, 32>(a, st); break;
    default: launch_nb<TB, VW, 64>(a, st); break;
    }
}

}  // namespace

dspmm_status_t launch_nnz_tiled(const NbLaunch &a, hipStream_t st) {
    if (a.ntiles == 0) return DSPMM_STATUS_SUCCESS;
    if (a.bf16) {
        switch (a.vw) {
        case 8: launch_nb_l<uint16_t, 8>(a, st); break;
        case 4: launch_nb_l<uint16_t, 4>(a, st); break;
        case 2: launch_nb_l<uint16_t, 2>(a, st); break;
        default: launch_nb_l<uint16_t, 1>(a, st); break;
        }
    } else {
        switch (a.vw) {
        case 4: launch_nb_l<float, 4>(a, st); break;
        case 2: launch_nb_l<float, 2>(a, st); break;
        default: launch_nb_l<float, 1>(a, st); break;
        }
    }
    return launch_status();
}

dspmm_status_t launch_fill_zero(void *p, size_t bytes, hipStream_t st) {
    if (bytes == 0) return DSPMM_STATUS_SUCCESS;
    if ((bytes % 4) == 0 && aligned(p, 4)) {
        const int64_t n = (int64_t)(bytes / 4);
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 256 * 32);
        hipLaunchKernelGGL(k_fill_zero_u32, dim3(blocks), dim3(256), 0, st,
                           reinterpret_cast<uint32_t *>(p), n);
    } else {
        const int64_t n = (int64_t)bytes;
        hipLaunchKernelGGL(k_fill_zero_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<uint8_t *>(p), n);
    }
    return launch_status();
}

dspmm_status_t launch_f32_to_bf16(const float *src, uint16_t *dst, int64_t n, hipStream_t st) {
    if (n == 0) return DSPMM_STATUS_SUCCESS;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks), dim3(256), 0, st, src, dst, n);
    return launch_status();
}

}  // namespace dspmm
