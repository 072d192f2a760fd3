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
// One workgroup per split row.  Canonical order, a function of (nv, nchunks)
// only: Lc = min(256, next power of two >= nv) threads cover the columns,
// the chunks are cut into P = 256 / Lc contiguous parts, each part is summed
// in chunk order by one thread per column, and the P part sums are added in
// part order.  Single-matrix and batched launches share this function, so
// their results agree bit for bit.  A thread loads up to 16 chunk values
// before adding them, so its loads overlap.
__device__ __forceinline__ void combine_split(const ItemDesc &d, int4 sp, float *lds) {
    const int row = sp.y, slot0 = sp.z, nch = sp.w;
    const int t = threadIdx.x;
    int Lc = 1;
    while (Lc < d.nv && Lc < kBlock) Lc <<= 1;
    const int P = kBlock / Lc;
    const int q = t / Lc, jl = t % Lc;
    const int per = (nch + P - 1) / P;
    const int c0 = min(nch, q * per), c1 = min(nch, c0 + per);
    for (int jb = 0; jb < d.nv; jb += Lc) {          // uniform loop
        const int j = jb + jl;
        float s = 0.f;
        if (j < d.nv) {
            const float *src = d.carry + (int64_t)slot0 * d.ldcarry + j;
            for (int c = c0; c < c1; c += 16) {
                float v[16];
#pragma unroll
                for (int u = 0; u < 16; u++) v[u] = (c + u < c1) ? src[(int64_t)(c + u) * d.ldcarry] : 0.f;
#pragma unroll
                for (int u = 0; u < 16; u++)
                    if (c + u < c1) s = s + v[u];
            }
        }
        if (P > 1) {
            lds[t] = s;
            __syncthreads();
            if (q == 0) {
                for (int qq = 1; qq < P; qq++) s = s + lds[qq * Lc + jl];
            }
            __syncthreads();
        }
        if (q == 0 && j < d.nv) d.C[(int64_t)row * d.ldc + j] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_split_combine(ItemDesc d, const int4 *__restrict__ splits,
                                                          int nsplits) {
    __shared__ float lds[kBlock];
    combine_split(d, splits[blockIdx.x], lds);
}

// Batched: workgroup b finds its item by binary search in the prefix sums
// of the items' split counts.
__global__ __launch_bounds__(kBlock) void k_split_combine_batched(const ItemDesc *__restrict__ items,
                                                                  const int *__restrict__ prefix,
                                                                  int count) {
    __shared__ float lds[kBlock];
    const int b = blockIdx.x;
    int lo = 0, hi = count - 1;                     // last item with prefix[i] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const ItemDesc &d = items[lo];
    combine_split(d, d.splits[b - prefix[lo]], lds);
}

// ---- SpMV (nv == 1) -----------------------------------------------------------
// One workgroup per tile of at most T nonzeros.  Phase 1 computes every
// product a_p * x[col_p] into LDS with coalesced loads of (col, val) and
// independent gathers of x.  Phase 2: in a tile of whole rows no longer
// than kSpmvRowMax, thread i sums row r0+i from LDS in nonzero order --
// bit-exact with the host loop; a tile holding (part of) one longer row is
// reduced by the workgroup in a fixed tree order and written to y (one
// chunk) or to its carry slot (several chunks, combined afterwards).
template <typename TB>
__global__ __launch_bounds__(kBlock) void k_spmv_tiled(ItemDesc d, int ntiles) {
    extern __shared__ float prod[];
    __shared__ float red[kBlock];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int4 tile = d.tiles[b];
    const int pb = tile.z, pe = tile.w, n = pe - pb;
    const int t = threadIdx.x;
    const TB *__restrict__ x = reinterpret_cast<const TB *>(d.B);
    for (int i = t; i < n; i += kBlock) {
        const int c = ld_stream(d.colIdx + pb + i);
        const float a = ld_stream(d.values + pb + i);
        float xv[1];
        load_b<TB, 1>(xv, x + (int64_t)c * d.ldb);
        prod[i] = a * xv[0];
    }
    __syncthreads();
    const bool tree = tile.y < 0 || (tile.y - tile.x == 1 && n > kSpmvRowMax);
    if (!tree) {
        const int r = tile.x + t;
        if (r < tile.y) {
            const int s0 = d.rowPtr[r] - pb, s1 = d.rowPtr[r + 1] - pb;
            float s = 0.f;
            for (int p = s0; p < s1; p++) s = s + prod[p];
            d.C[(int64_t)r * d.ldc] = s;
        }
        return;
    }
    // fixed-order tree: thread t sums t, t+256, ... then pairwise halving
    float s = 0.f;
    for (int i = t; i < n; i += kBlock) s = s + prod[i];
    red[t] = s;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (t < h) red[t] = red[t] + red[t + h];
        __syncthreads();
    }
    if (t == 0) {
        if (tile.y < 0) d.carry[(int64_t)(-1 - tile.y) * d.ldcarry] = red[0];
        else d.C[(int64_t)tile.x * d.ldc] = red[0];
    }
}

__global__ void k_fill_zero_u32(uint32_t *__restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}
__global__ void k_fill_zero_u8(uint8_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Seeded U[lo, hi) fill (SplitMix64 of seed ^ index, 24-bit mantissa);
// BF16 output is the RNE rounding of the same FP32 value.
__device__ __forceinline__ float seeded_uniform(uint64_t seed, int64_t i, float lo, float hi) {
    uint64_t z = seed ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = (float)(z >> 40) * 0x1.0p-24f;
    return lo + (hi - lo) * u;
}
__device__ __forceinline__ uint16_t to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) u |= 0x00400000u;
    else u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__global__ void k_fill_uniform(void *__restrict__ out, int64_t n, uint64_t seed, int bf16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = seeded_uniform(seed, i, -1.f, 1.f);
        if (bf16) reinterpret_cast<uint16_t *>(out)[i] = to_bf16_rne(v);
        else reinterpret_cast<float *>(out)[i] = v;
    }
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
        hipLaunchKernelGGL(k_split_combine, dim3(a.nsplits), dim3(kBlock), 0, st, a.item, a.splits,
                           a.nsplits);
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
        hipLaunchKernelGGL(k_split_combine_batched, dim3(a.nsplits), dim3(kBlock), 0, st, a.items,
                           a.split_prefix, a.count);
        return launch_status();
    }
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t launch_spmv_tiled(const NbLaunch &a, hipStream_t st) {
    if (a.ntiles == 0) return DSPMM_STATUS_SUCCESS;
    const size_t lds = sizeof(float) * (size_t)a.item.S;
    if (a.bf16)
        hipLaunchKernelGGL(k_spmv_tiled<uint16_t>, dim3(a.ntiles), dim3(kBlock), lds, st, a.item, a.ntiles);
    else
        hipLaunchKernelGGL(k_spmv_tiled<float>, dim3(a.ntiles), dim3(kBlock), lds, st, a.item, a.ntiles);
    dspmm_status_t s = launch_status();
    if (s != DSPMM_STATUS_SUCCESS) return s;
    if (a.mid_event) (void)hipEventRecord(a.mid_event, st);
    if (a.nsplits > 0) {
        hipLaunchKernelGGL(k_split_combine, dim3(a.nsplits), dim3(kBlock), 0, st, a.item, a.splits,
                           a.nsplits);
        s = launch_status();
    }
    return s;
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

dspmm_status_t launch_fill_uniform(void *out, int64_t n, uint64_t seed, bool bf16, hipStream_t st) {
    if (n == 0) return DSPMM_STATUS_SUCCESS;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_fill_uniform, dim3(blocks), dim3(256), 0, st, out, n, seed, bf16 ? 1 : 0);
    return launch_status();
}

dspmm_status_t launch_f32_to_bf16(const float *src, uint16_t *dst, int64_t n, hipStream_t st) {
    if (n == 0) return DSPMM_STATUS_SUCCESS;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks), dim3(256), 0, st, src, dst, n);
    return launch_status();
}

}  // namespace dspmm
