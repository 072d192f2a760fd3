// native_kernels.hip -- the native load-balanced SpMM path: single-matrix
// and batched (persistent) nnz-tiled kernels, the split-row combine, and
// utility kernels.
//
// Nnz-tiled SpMM (row-major, overwrite, deterministic).  A host-built plan
// cuts the matrix into tiles of at most S nonzeros that start and end on
// row boundaries; a row longer than S is cut into S-sized chunks whose
// partial sums go to a carry buffer and are combined, in one fixed order,
// by a second small kernel.  Every tile is owned by a group of L lanes x VW
// floats that sums each of its rows sequentially, in nonzero order, with a
// separately rounded multiply and add -- so every row that is not split
// equals the reference host loop (spmvspmm_benchmark.cu:11-39) bit for
// bit, with FP32 or BF16-stored B.  A wavefront owns up to G = 64/L
// consecutive tiles; it stages their (col, val) and row offsets through LDS
// with coalesced loads, then issues UNR B-row gathers per group before
// summing them.  Stores of C are non-temporal so C does not evict B from
// L2 / MALL.
//
// The batched kernel runs the same per-wave code over the tiles of many
// matrices in one launch: each XCD owns a list of (matrix, unit range)
// segments, the workgroups resident on that XCD walk it, and a matrix whose
// B fits in L2 is kept on one XCD.  Results are bit-identical to the
// single-matrix kernel for the same tile size.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "common.hpp"
#include "internal.hpp"
#include "native.hpp"

namespace dspmm {

// LDS position of staged element i: one pad word every 16 elements, so that
// groups reading their tiles' k-th element at 16-element strides spread
// over banks.
__device__ __forceinline__ int skew(int i) { return i + (i >> 4); }
__host__ __device__ inline int skew_h(int i) { return i + (i >> 4); }

// Ints of LDS one wavefront needs for G tiles of at most S nonzeros / rows.
__host__ __device__ inline int wave_lds_ints(int G, int S) {
    return 2 * skew_h(G * S) + G * S + 1;
}

struct WaveStage {
    int P0, R0;
    int *s_col;
    float *s_val;
    int *s_rp;
};

// Stages the (col, val) of tiles [t0, t0+tcount) of item d and the row ends
// of the rows they cover into the wave's LDS region (whole wavefront).
__device__ __forceinline__ WaveStage stage_tiles(const ItemDesc &d, int t0, int tcount, int G,
                                                 int *lds) {
    WaveStage w;
    const int cap = G * d.S;
    w.s_col = lds;
    w.s_val = reinterpret_cast<float *>(lds + skew(cap));
    w.s_rp = lds + 2 * skew(cap);
    w.P0 = 0;
    w.R0 = 0;
    if (tcount <= 0) return w;
    const int ln = lane_id();
    const int4 tf = d.tiles[t0];
    const int4 tl = d.tiles[t0 + tcount - 1];
    w.P0 = tf.z;
    w.R0 = tf.x;
    const int P1 = tl.w;
    const int R1 = tl.y >= 0 ? tl.y : tl.x + 1;
    for (int i = ln; i < P1 - w.P0; i += kWave) {
        w.s_col[skew(i)] = ld_stream(d.colIdx + w.P0 + i);
        w.s_val[skew(i)] = ld_stream(d.values + w.P0 + i);
    }
    for (int i = ln; i < R1 - w.R0; i += kWave) w.s_rp[i] = d.rowPtr[w.R0 + 1 + i];
    return w;
}

// Group g of the wavefront processes tile t0+g (if g < tcount) on column
// slab `slab`.
template <typename TB, int VW, int L, int UNR>
__device__ __forceinline__ void process_tiles(const ItemDesc &d, int t0, int tcount, int slab,
                                              const WaveStage &w) {
    const int ln = lane_id();
    const int g = ln / L;
    const int col0 = slab * d.slab_w + (ln % L) * VW;
    if (g >= tcount || col0 >= d.nv) return;
    const TB *__restrict__ B = reinterpret_cast<const TB *>(d.B);
    const int4 tile = d.tiles[t0 + g];
    const int pb = tile.z, pe = tile.w;
    const int P0 = w.P0, R0 = w.R0;
    const int *s_col = w.s_col;
    const float *s_val = w.s_val;
    float acc[VW];
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = 0.f;

    if (tile.y < 0) {
        // Chunk of a split row: sum it and leave it in its carry slot.
        for (int p = pb; p < pe; p += UNR) {
            float bv[UNR][VW];
#pragma unroll
            for (int u = 0; u < UNR; u++)
                if (p + u < pe)
                    load_b<TB, VW>(bv[u], B + (int64_t)s_col[skew(p + u - P0)] * d.ldb + col0);
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
        store_vec<VW>(d.carry + (int64_t)slot * d.ldcarry + col0, acc);
        return;
    }

    int r = tile.x;
    const int re = tile.y;
    int next_end = w.s_rp[r - R0];                  // rowPtr[r + 1]
    float *__restrict__ C = d.C;
    for (int p = pb; p < pe; p += UNR) {
        float bv[UNR][VW];
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (p + u < pe)
                load_b<TB, VW>(bv[u], B + (int64_t)s_col[skew(p + u - P0)] * d.ldb + col0);
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int q = p + u;
            if (q < pe) {
                while (q >= next_end) {           // row r is complete (or empty)
                    store_vec_nt<VW>(C + (int64_t)r * d.ldc + col0, acc);
#pragma unroll
                    for (int k = 0; k < VW; k++) acc[k] = 0.f;
                    ++r;
                    next_end = w.s_rp[r - R0];
                }
                const float a = s_val[skew(q - P0)];
#pragma unroll
                for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * bv[u][k];
            }
        }
    }
    for (; r < re; ++r) {                          // last row, then trailing empty rows
        store_vec_nt<VW>(C + (int64_t)r * d.ldc + col0, acc);
#pragma unroll
        for (int k = 0; k < VW; k++) acc[k] = 0.f;
    }
}

// Single matrix: wavefront w of the grid owns tiles [w*G, w*G+G).
template <typename TB, int VW, int L, int UNR>
__global__ __launch_bounds__(kBlock) void k_nnz_tiled(ItemDesc d, int ntiles) {
    constexpr int G = kWave / L;
    extern __shared__ int smem[];
    const int wib = (int)threadIdx.x / kWave;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int t0 = (b * (kBlock / kWave) + wib) * G;
    const int tcount = t0 < ntiles ? min(G, ntiles - t0) : 0;
    const WaveStage w = stage_tiles(d, t0, tcount, G, smem + wib * wave_lds_ints(G, d.S));
    __syncthreads();
    process_tiles<TB, VW, L, UNR>(d, t0, tcount, blockIdx.y, w);
}

// ---- batched -------------------------------------------------------------
// Class of an item = (VW, L).  The kernel dispatches per work unit.
template <typename TB, typename F>
__device__ __forceinline__ void with_class(int cls, F &&f) {
    // cls = 8 * log2(VW) + log2(L)
    switch (cls) {
#define DSPMM_CLS(VW_, LG_)                                                              \
    case 8 * ((VW_) == 1 ? 0 : (VW_) == 2 ? 1 : (VW_) == 4 ? 2 : 3) + (LG_):           \
        f(std::integral_constant<int, VW_>{}, std::integral_constant<int, 1 << (LG_)>{}); \
        break;
#define DSPMM_CLS_ALL_L(VW_) \
    DSPMM_CLS(VW_, 0) DSPMM_CLS(VW_, 1) DSPMM_CLS(VW_, 2) DSPMM_CLS(VW_, 3) DSPMM_CLS(VW_, 4) \
    DSPMM_CLS(VW_, 5) DSPMM_CLS(VW_, 6)
        DSPMM_CLS_ALL_L(1)
        DSPMM_CLS_ALL_L(2)
        DSPMM_CLS_ALL_L(4)
#undef DSPMM_CLS_ALL_L
#undef DSPMM_CLS
    default:
        if constexpr (sizeof(TB) == 2) {
            switch (cls - 24) {
            case 0: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 1>{}); break;
            case 1: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 2>{}); break;
            case 2: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 4>{}); break;
            case 3: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 8>{}); break;
            case 4: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 16>{}); break;
            case 5: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 32>{}); break;
            default: f(std::integral_constant<int, 8>{}, std::integral_constant<int, 64>{}); break;
            }
        }
        break;
    }
}

// Persistent batched kernel.  segs[xoff[x] .. xoff[x+1]) are XCD x's work
// segments {item, first unit of the item, unit count, units before this
// segment in x's list}; xoff[8+x] is x's total unit count.  Workgroups
// resident on XCD x (blockIdx.x % 8 == x) take consecutive groups of four
// units of x's list, one per wavefront, until the list is exhausted; every
// wavefront of a workgroup runs the same number of iterations, so the
// workgroup barrier that separates LDS staging from its use is uniform.
template <typename TB>
__global__ __launch_bounds__(kBlock) void k_nnz_tiled_batched(const ItemDesc *__restrict__ items,
                                                              const int4 *__restrict__ segs,
                                                              const int *__restrict__ xoff,
                                                              int lds_ints_per_wave) {
    extern __shared__ int smem[];
    const int wib = (int)threadIdx.x / kWave;
    const int x = blockIdx.x % kNumXcd;
    const int bx = blockIdx.x / kNumXcd;
    const int nbx = ((int)gridDim.x - x + kNumXcd - 1) / kNumXcd;
    const int s_lo = xoff[x], s_hi = xoff[x + 1];
    const int nunits = xoff[kNumXcd + 1 + x];
    int *lds = smem + wib * lds_ints_per_wave;
    for (int k = bx; k * (kBlock / kWave) < nunits; k += nbx) {
        const int u = k * (kBlock / kWave) + wib;
        int item = -1, lu = 0;
        if (u < nunits) {
            // last segment whose start <= u (segments sorted by .w)
            int lo = s_lo, hi = s_hi - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (segs[mid].w <= u) lo = mid; else hi = mid - 1;
            }
            const int4 sg = segs[lo];
            item = sg.x;
            lu = sg.y + (u - sg.w);
        }
        ItemDesc d;
        int t0 = 0, tcount = 0, slab = 0, G = 1;
        if (item >= 0) {
            d = items[item];
            G = kWave >> (d.cls & 7);
            const int per_slab = (d.ntiles + G - 1) / G;
            slab = lu / per_slab;
            t0 = (lu % per_slab) * G;
            tcount = min(G, d.ntiles - t0);
        }
        WaveStage w{};
        if (item >= 0) w = stage_tiles(d, t0, tcount, G, lds);
        __syncthreads();
        if (item >= 0) {
            with_class<TB>(d.cls, [&](auto vw_c, auto l_c) {
                constexpr int VW_ = decltype(vw_c)::value, L_ = decltype(l_c)::value;
                if constexpr (sizeof(TB) == 4 && VW_ == 8) {
                } else {
                    process_tiles<TB, VW_, L_, 8>(d, t0, tcount, slab, w);
                }
            });
        }
        __syncthreads();
    }
}

// ---- split-row combine ------------------------------------------------------
// One wavefront per split row.  For every column j the chunk partial sums
// are added in one canonical order: lane l sums chunks l, l+64, l+128, ...
// in sequence, then the 64 lane sums are combined by a butterfly (xor 1, 2,
// 4, ..., 32; every lane ends with the same value).  Single-matrix and
// batched launches use this same function, so their results agree bit for
// bit.
__device__ __forceinline__ void combine_split(const ItemDesc &d, int4 sp) {
    const int row = sp.y, slot0 = sp.z, nch = sp.w;
    const int ln = lane_id();
    for (int j0 = 0; j0 < d.nv; j0 += 4) {
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = 0.f;
        for (int c = ln; c < nch; c += kWave) {
            const float *src = d.carry + (int64_t)(slot0 + c) * d.ldcarry + j0;
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = v[k] + src[k];   // ldcarry pads to 8, j0+k < ldcarry
        }
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = butterfly_sum<1, kWave>(v[k]);
        if (ln < 4 && j0 + ln < d.nv) {
            float o = v[0];
            if (ln == 1) o = v[1];
            if (ln == 2) o = v[2];
            if (ln == 3) o = v[3];
            d.C[(int64_t)row * d.ldc + j0 + ln] = o;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_split_combine(ItemDesc d, const int4 *__restrict__ splits,
                                                          int nsplits) {
    const int w = (int)(blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (w >= nsplits) return;                        // wave-uniform
    combine_split(d, splits[w]);
}

__global__ __launch_bounds__(kBlock) void k_split_combine_batched(const ItemDesc *__restrict__ items,
                                                                  const int4 *__restrict__ splits,
                                                                  int nsplits) {
    const int w = (int)(blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (w >= nsplits) return;                        // wave-uniform
    const int4 sp = splits[w];
    combine_split(items[sp.x], sp);
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
int item_class(int vw, int lanes) {
    int lg = 0;
    while ((1 << lg) < lanes) lg++;
    const int vc = vw == 1 ? 0 : vw == 2 ? 1 : vw == 4 ? 2 : 3;
    return 8 * vc + lg;
}

size_t nnz_tiled_lds_bytes(int lanes, int S) {
    return (size_t)(kBlock / kWave) * (size_t)wave_lds_ints(kWave / lanes, S) * sizeof(int);
}

namespace {

template <typename TB, int VW, int L>
void launch_nb(const NbLaunch &a, hipStream_t st) {
    constexpr int G = kWave / L;
    const int waves = div_up(a.ntiles, G);
    dim3 grid(div_up(waves, kBlock / kWave), a.nslabs);
    const size_t lds = nnz_tiled_lds_bytes(L, a.item.S);
    if (a.unroll <= 4)
        hipLaunchKernelGGL((k_nnz_tiled<TB, VW, L, 4>), grid, dim3(kBlock), lds, st, a.item, a.ntiles);
    else
        hipLaunchKernelGGL((k_nnz_tiled<TB, VW, L, 8>), grid, dim3(kBlock), lds, st, a.item, a.ntiles);
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
    dspmm_status_t s = launch_status();
    if (s != DSPMM_STATUS_SUCCESS) return s;
    if (a.mid_event) (void)hipEventRecord(a.mid_event, st);
    if (a.nsplits > 0) {
        hipLaunchKernelGGL(k_split_combine, dim3(div_up(a.nsplits, kBlock / kWave)), dim3(kBlock), 0,
                           st, a.item, a.splits, a.nsplits);
        s = launch_status();
    }
    return s;
}

dspmm_status_t launch_nnz_tiled_batched(const BatchLaunch &a, hipStream_t st) {
    if (a.nblocks > 0) {
        const size_t lds = (size_t)(kBlock / kWave) * (size_t)a.lds_ints_per_wave * sizeof(int);
        if (a.bf16)
            hipLaunchKernelGGL(k_nnz_tiled_batched<uint16_t>, dim3(a.nblocks), dim3(kBlock), lds, st,
                               a.items, a.segs, a.xoff, a.lds_ints_per_wave);
        else
            hipLaunchKernelGGL(k_nnz_tiled_batched<float>, dim3(a.nblocks), dim3(kBlock), lds, st,
                               a.items, a.segs, a.xoff, a.lds_ints_per_wave);
        dspmm_status_t s = launch_status();
        if (s != DSPMM_STATUS_SUCCESS) return s;
    }
    if (a.mid_event) (void)hipEventRecord(a.mid_event, st);
    if (a.nsplits > 0) {
        hipLaunchKernelGGL(k_split_combine_batched, dim3(div_up(a.nsplits, kBlock / kWave)),
                           dim3(kBlock), 0, st, a.items, a.splits, a.nsplits);
        return launch_status();
    }
    return DSPMM_STATUS_SUCCESS;
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
