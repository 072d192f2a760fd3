// native_kernels.hip -- the native SpMM path: single-matrix and batched
// (persistent) nnz-tiled kernels, the split-row combine, the SpMV tiles,
// the column-cached, row-blocked and row-sequential kernels, and utility
// kernels (fills, BF16 conversion, transposes).
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

// Wavefront-level LDS ordering: every wavefront stages into and reads from
// its own LDS region only, so a workgroup barrier is not needed between
// staging and use -- the wavefront's LDS writes are made complete and
// ordered before its reads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
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
    wave_lds_sync();
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
// segment in x's list}, in item order; xoff[8+1+x] is x's total unit count.
// The workgroups resident on XCD x (blockIdx.x % 8 == x) take x's units in
// list order, four consecutive units (one per wavefront) per workgroup and
// step, so at any moment an XCD works on a narrow window of its list (few
// matrices, whose B rows then stay in that XCD's L2).  A wavefront finds its
// first unit's segment by binary search and then walks forward; it stages
// into and reads from its own LDS region only, so wavefronts never wait for
// one another.
template <typename TB>
__global__ __launch_bounds__(kBlock) void k_nnz_tiled_batched(const ItemDesc *__restrict__ items,
                                                              const int4 *__restrict__ segs,
                                                              const int *__restrict__ xoff,
                                                              int lds_ints_per_wave) {
    extern __shared__ int smem[];
    constexpr int W = kBlock / kWave;
    const int wib = (int)threadIdx.x / kWave;
    const int x = blockIdx.x % kNumXcd;
    const int bx = blockIdx.x / kNumXcd;
    const int nbx = ((int)gridDim.x - x + kNumXcd - 1) / kNumXcd;
    const int s_lo = xoff[x], s_hi = xoff[x + 1];
    const int nunits = xoff[kNumXcd + 1 + x];
    int *lds = smem + wib * lds_ints_per_wave;
    int seg = -1, seg_end = 0;
    int4 sg = make_int4(0, 0, 0, 0);
    for (int u = bx * W + wib; u < nunits; u += nbx * W) {
        if (seg < 0) {                              // last segment whose start <= u
            int lo = s_lo, hi = s_hi - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (segs[mid].w <= u) lo = mid; else hi = mid - 1;
            }
            seg = lo;
            sg = segs[seg];
            seg_end = sg.w + sg.z;
        }
        while (u >= seg_end) {                      // walk forward (units only increase)
            sg = segs[++seg];
            seg_end = sg.w + sg.z;
        }
        const int item = sg.x, lu = sg.y + (u - sg.w);
        const ItemDesc d = items[item];
        const int G = kWave >> (d.cls & 7);
        const int per_slab = (d.ntiles + G - 1) / G;
        const int slab = lu / per_slab;
        const int t0 = (lu % per_slab) * G;
        const int tcount = min(G, d.ntiles - t0);
        const WaveStage w = stage_tiles(d, t0, tcount, G, lds);
        wave_lds_sync();
        with_class<TB>(d.cls, [&](auto vw_c, auto l_c) {
            constexpr int VW_ = decltype(vw_c)::value, L_ = decltype(l_c)::value;
            if constexpr (sizeof(TB) == 4 && VW_ == 8) {
            } else {
                process_tiles<TB, VW_, L_, 8>(d, t0, tcount, slab, w);
            }
        });
        wave_lds_sync();                            // LDS is restaged for the next unit
    }
}

// ---- split-row combine ------------------------------------------------------
// One wavefront per split row, four per workgroup.  Canonical order, a
// function of (nv, nchunks) only: Lw = min(64, next power of two >= nv)
// lanes cover a block of columns (column blocks of Lw in turn when nv > 64),
// the chunks are cut into P = 64 / Lw contiguous parts, each part is summed
// in chunk order by one lane per column, and the first part's lanes add the
// P part sums in part order (read across lanes with ds_bpermute).
// Single-matrix and batched launches share this function, so their results
// agree bit for bit.  A lane loads up to 16 chunk values before adding them,
// so its loads overlap.
// Column blocks cb0, cb0 + cstep, ... of the row (cstep = 1: all of them).
// A column's sum does not depend on which wavefront computes its block.
__device__ __forceinline__ int combine_lanes(int nv) {
    int Lw = 1;
    while (Lw < nv && Lw < kWave) Lw <<= 1;
    return Lw;
}
__device__ __forceinline__ void combine_split(const ItemDesc &d, int4 sp, int cb0, int cstep) {
    const int row = sp.y, slot0 = sp.z, nch = sp.w;
    const int ln = lane_id();
    const int Lw = combine_lanes(d.nv);
    const int P = kWave / Lw;
    const int q = ln / Lw, jl = ln % Lw;
    const int per = (nch + P - 1) / P;
    const int c0 = min(nch, q * per), c1 = min(nch, c0 + per);
    for (int jb = cb0 * Lw; jb < d.nv; jb += cstep * Lw) {   // wave-uniform loop
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
        for (int qq = 1; qq < P; qq++) {             // wave-uniform; every lane takes part
            const float o = __shfl(s, qq * Lw + jl, kWave);
            if (q == 0) s = s + o;
        }
        if (q == 0 && j < d.nv) d.C[(int64_t)row * d.ldc + j] = s;
    }
}

// One wavefront per (split row, column block of combine_lanes(nv) columns):
// rows with nv > 64 spread their column blocks over wavefronts, which
// shortens the longest rows' chains without changing any sum.
// Workgroups of k_split_combine: one wavefront per (split row, column block).
int combine_blocks(int nsplits, int nv) {
    int lw = 1;
    while (lw < nv && lw < kWave) lw <<= 1;
    return (int)div_up((int64_t)nsplits * div_up(nv, lw), (int64_t)(kBlock / kWave));
}

__global__ __launch_bounds__(kBlock) void k_split_combine(ItemDesc d, const int4 *__restrict__ splits,
                                                          int nsplits) {
    const int nblk = (d.nv + combine_lanes(d.nv) - 1) / combine_lanes(d.nv);
    const int64_t w = (int64_t)blockIdx.x * (kBlock / kWave) + (int)threadIdx.x / kWave;
    if (w >= (int64_t)nsplits * nblk) return;       // whole wavefronts leave; no barrier follows
    combine_split(d, splits[w / nblk], (int)(w % nblk), nblk);
}

// Batched: wavefront w finds its item by binary search in the prefix sums
// of the items' split counts.
__global__ __launch_bounds__(kBlock) void k_split_combine_batched(const ItemDesc *__restrict__ items,
                                                                  const int *__restrict__ prefix,
                                                                  int count, int nsplits) {
    const int w = blockIdx.x * (kBlock / kWave) + (int)threadIdx.x / kWave;
    if (w >= nsplits) return;
    int lo = 0, hi = count - 1;                     // last item with prefix[i] <= w
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= w) lo = mid; else hi = mid - 1;
    }
    const ItemDesc &d = items[lo];
    combine_split(d, d.splits[w - prefix[lo]], 0, 1);
}

// ---- SpMV (nv == 1) -----------------------------------------------------------
// One workgroup per tile.  A tile of whole rows (each no longer than
// kSpmvRowMax, at most T = E * 256 nonzeros): every thread first issues its
// E (col, val) loads and its row bounds, then E independent gathers of x,
// and writes the products to LDS.  After one barrier, thread i sums row
// r0+i from LDS in nonzero order if the row has at most kSpmvThreadRow
// nonzeros -- bit-exact with the host loop; longer rows of the tile are
// queued in LDS and taken one per wavefront: lane l sums products l, l+64,
// ... in order, then a fixed butterfly.  A tile of one longer row (up to
// kSpmvLongTiles * T nonzeros) or of one T-chunk of a still longer row is
// reduced by the workgroup: thread t sums products t, t+256, t+512, ... in
// that order, then a pairwise tree; the result goes to y (whole row) or to
// the chunk's carry slot.
template <typename TB, int E>
__global__ __launch_bounds__(kBlock) void k_spmv_tiled(ItemDesc d, int ntiles) {
    __shared__ float prod[E * kBlock];
    __shared__ float red[kBlock];
    const int4 tile = d.tiles[blockIdx.x];
    const int pb = tile.z, n = tile.w - tile.z;
    const int t = threadIdx.x;
    const TB *__restrict__ x = reinterpret_cast<const TB *>(d.B);
    const int *__restrict__ ci = d.colIdx + pb;
    const float *__restrict__ va = d.values + pb;
    const bool tree = tile.y < 0 || (tile.y - tile.x == 1 && n > kSpmvRowMax);
    if (!tree) {
        __shared__ int nlong;
        __shared__ int4 lrange[kBlock];
        if (t == 0) nlong = 0;
        const int r = tile.x + t;
        int s0 = 0, s1 = 0;
        if (r < tile.y) {
            s0 = d.rowPtr[r] - pb;
            s1 = d.rowPtr[r + 1] - pb;
        }
        int c[E];
        float a[E];
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int i = t + k * kBlock;
            c[k] = 0;
            a[k] = 0.f;
            if (i < n) {
                c[k] = ld_stream(ci + i);
                a[k] = ld_stream(va + i);
            }
        }
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int i = t + k * kBlock;
            if (i < n) {
                float xv[1];
                load_b<TB, 1>(xv, x + (int64_t)c[k] * d.ldb);
                prod[i] = a[k] * xv[0];
            }
        }
        __syncthreads();
        if (r < tile.y) {
            if (s1 - s0 <= kSpmvThreadRow) {
                float s = 0.f;
                for (int p = s0; p < s1; p++) s = s + prod[p];
                d.C[(int64_t)r * d.ldc] = s;
            } else {
                lrange[atomicAdd(&nlong, 1)] = make_int4(s0, s1, r, 0);
            }
        }
        __syncthreads();
        const int wv = t / kWave, ln = t % kWave;
        for (int k = wv; k < nlong; k += kBlock / kWave) {      // wave-uniform
            const int4 q = lrange[k];
            float s = 0.f;
            for (int p = q.x + ln; p < q.y; p += kWave) s = s + prod[p];
            s = butterfly_sum<1, kWave>(s);
            if (ln == 0) d.C[(int64_t)q.z * d.ldc] = s;
        }
        return;
    }
    float s = 0.f;
    int i = t;
    for (; i + 3 * kBlock < n; i += 4 * kBlock) {
        int c[4];
        float a[4], xv[4][1];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            c[k] = ld_stream(ci + i + k * kBlock);
            a[k] = ld_stream(va + i + k * kBlock);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) load_b<TB, 1>(xv[k], x + (int64_t)c[k] * d.ldb);
#pragma unroll
        for (int k = 0; k < 4; k++) s = s + a[k] * xv[k][0];
    }
    for (; i < n; i += kBlock) {
        float xv[1];
        load_b<TB, 1>(xv, x + (int64_t)ld_stream(ci + i) * d.ldb);
        s = s + ld_stream(va + i) * xv[0];
    }
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

// ---- column-cached SpMM (row-major) -------------------------------------------
// For matrices whose neighbouring rows share columns (stencils, banded
// matrices): a host plan cuts the rows into tiles of whole rows whose
// distinct columns number at most umax, and gives every nonzero the position
// of its column in its tile's column list.  One workgroup per tile copies the
// tile's B rows into LDS once (coalesced, 16 B per lane), stages the tile's
// (position, value) pairs and row ends, and then every row is summed by a
// group of L lanes x VW columns from LDS, in nonzero order with a separately
// rounded multiply and add -- bit-exact with the host loop.  Each B row is
// read from L2 once per tile instead of once per nonzero.  A row whose own
// columns exceed umax is a "direct" tile: the G groups take contiguous
// chunks of it, gather B from global memory, and group 0 adds the G
// partial rows in group order.
template <int VW, typename TB>
__device__ __forceinline__ void copy_raw(TB *dst, const TB *src) {
    constexpr int bytes = VW * (int)sizeof(TB);
    if constexpr (bytes == 16) *reinterpret_cast<uint4 *>(dst) = *reinterpret_cast<const uint4 *>(src);
    else if constexpr (bytes == 8) *reinterpret_cast<uint2 *>(dst) = *reinterpret_cast<const uint2 *>(src);
    else if constexpr (bytes == 4) *reinterpret_cast<uint32_t *>(dst) = *reinterpret_cast<const uint32_t *>(src);
    else *reinterpret_cast<uint16_t *>(dst) = *reinterpret_cast<const uint16_t *>(src);
}

template <typename TB, int VW>
__device__ __forceinline__ void widen(float (&r)[VW], const TB *p) {
    if constexpr (sizeof(TB) == 4) {
#pragma unroll
        for (int k = 0; k < VW; k++) r[k] = reinterpret_cast<const float *>(p)[k];
    } else {
#pragma unroll
        for (int k = 0; k < VW; k++)
            r[k] = __uint_as_float((uint32_t)reinterpret_cast<const uint16_t *>(p)[k] << 16);
    }
}

// LDS row stride of the cached B rows: nv rounded up to VW, plus VW padding
// elements so that groups reading different rows start on different banks.
__host__ __device__ inline int cached_stride(int nv, int vw) { return (nv + vw - 1) / vw * vw + vw; }

template <typename TB, int VW, int L, int NT>
__global__ __launch_bounds__(NT) void k_nnz_cached(CachedLaunch a) {
    constexpr int G = NT / L;
    extern __shared__ __align__(16) unsigned char csm[];
    const int t = threadIdx.x;
    const int g = t / L;
    const int col0 = (t % L) * VW;
    const int ls = cached_stride(a.nv, VW);
    TB *Bl = reinterpret_cast<TB *>(csm);
    float *lv = reinterpret_cast<float *>(csm + (size_t)a.umax * ls * sizeof(TB));
    uint16_t *li = reinterpret_cast<uint16_t *>(lv + a.smax);
    int *rend = reinterpret_cast<int *>(li + ((a.smax + 1) & ~1));
    const int ti = xcd_remap(blockIdx.x, gridDim.x);
    const int4 tile = a.tiles[ti];
    const TB *__restrict__ B = reinterpret_cast<const TB *>(a.B);

    if (tile.y < 0) {                                  // direct tile: one long row
        const int row = tile.x, p0 = tile.z, p1 = a.rowPtr[row + 1];
        const int len = p1 - p0, ch = (len + G - 1) / G;
        const int c0 = p0 + min(len, g * ch), c1 = p0 + min(len, (g + 1) * ch);
        float acc[VW];
#pragma unroll
        for (int k = 0; k < VW; k++) acc[k] = 0.f;
        if (col0 < a.nv) {
            for (int p = c0; p < c1; p++) {
                float bv[VW];
                load_b<TB, VW>(bv, B + (int64_t)ld_stream(a.colIdx + p) * a.ldb + col0);
                const float v = ld_stream(a.values + p);
#pragma unroll
                for (int k = 0; k < VW; k++) acc[k] = acc[k] + v * bv[k];
            }
        }
        float *part = reinterpret_cast<float *>(csm);   // G partial rows of stride ls
        if (col0 < a.nv) {
#pragma unroll
            for (int k = 0; k < VW; k++) part[g * ls + col0 + k] = acc[k];
        }
        __syncthreads();
        if (g == 0 && col0 < a.nv) {
            float s[VW];
#pragma unroll
            for (int k = 0; k < VW; k++) s[k] = part[col0 + k];
            for (int q = 1; q < G; q++) {
#pragma unroll
                for (int k = 0; k < VW; k++) s[k] = s[k] + part[q * ls + col0 + k];
            }
            store_vec_nt<VW>(a.C + (int64_t)row * a.ldc + col0, s);
        }
        return;
    }

    const int r0 = tile.x, r1 = tile.y, p0 = tile.z;
    const int u0 = tile.w, nu = a.tiles[ti + 1].w - u0;
    const int np = a.rowPtr[r1] - p0;
    // stage the tile's B rows (group g copies rows g, g+G, ...)
    for (int u = g; u < nu; u += G) {
        if (col0 < a.nv) copy_raw<VW>(Bl + u * ls + col0, B + (int64_t)a.ucols[u0 + u] * a.ldb + col0);
    }
    for (int i = t; i < np; i += NT) {
        lv[i] = ld_stream(a.values + p0 + i);
        li[i] = a.lidx[p0 + i];
    }
    for (int i = t; i < r1 - r0; i += NT) rend[i] = a.rowPtr[r0 + 1 + i] - p0;
    __syncthreads();
    if (col0 >= a.nv) return;
    for (int rr = g; rr < r1 - r0; rr += G) {
        const int s0 = rr == 0 ? 0 : rend[rr - 1], s1 = rend[rr];
        float acc[VW];
#pragma unroll
        for (int k = 0; k < VW; k++) acc[k] = 0.f;
        int p = s0;
        for (; p + 3 < s1; p += 4) {
            float bv[4][VW];
#pragma unroll
            for (int u = 0; u < 4; u++) widen<TB, VW>(bv[u], Bl + (int)li[p + u] * ls + col0);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float v = lv[p + u];
#pragma unroll
                for (int k = 0; k < VW; k++) acc[k] = acc[k] + v * bv[u][k];
            }
        }
        for (; p < s1; p++) {
            float bv[VW];
            widen<TB, VW>(bv, Bl + (int)li[p] * ls + col0);
            const float v = lv[p];
#pragma unroll
            for (int k = 0; k < VW; k++) acc[k] = acc[k] + v * bv[k];
        }
        store_vec_nt<VW>(a.C + (int64_t)(r0 + rr) * a.ldc + col0, acc);
    }
}

// ---- row-blocked SpMM (row-major) -------------------------------------------
// For matrices whose rows share columns with the rows a fixed distance
// away (stencils, banded matrices).  A host plan groups the rows in blocks
// of R rows `stride` apart -- block b holds rows
// (b / stride) * R * stride + b % stride + i * stride, i < R, so
// consecutive blocks hold neighbouring rows -- and lists, per block, the
// distinct columns of those rows in ascending order, each with a mask of
// the block's rows that have it ({col, mask}), and the block's values in
// that (column, row) order.  A group of L lanes x VW columns owns a block:
// it loads each listed B row once and adds it into the accumulators of the
// rows in its mask.  Rows are sorted, so each row still sums its nonzeros
// in nonzero order with a separately rounded multiply and add -- bit-exact
// with the host loop -- while the B rows a block's rows share are gathered
// once instead of once per row.  A 27-point stencil in blocks of R rows
// one grid plane apart lists 9 (R + 2) columns for 27 R nonzeros, and the
// next block (the neighbouring rows) finds two thirds of them in L1.
//
// A wavefront owns G = 64 / L consecutive row blocks.  Each group stages
// its own block's list entries and values into its slice of the
// wavefront's LDS region (kRbWin / G entries per window) with coalesced
// loads, so all G groups work in every window; the groups then read their
// (col, mask, values) from LDS as broadcasts.  The only per-entry
// vector-memory instruction left is the B-row gather, which is what the
// L1 address/data path is busy with.
constexpr int kRbWin = 128;

template <int R>
__device__ __forceinline__ void lds_row_vals(float (&v)[R], const float *p) {
    if constexpr (R >= 4) {
#pragma unroll
        for (int q = 0; q < R / 4; q++) {
            const float4 a = reinterpret_cast<const float4 *>(p)[q];
            v[4 * q] = a.x; v[4 * q + 1] = a.y; v[4 * q + 2] = a.z; v[4 * q + 3] = a.w;
        }
    } else {
        const float2 a = *reinterpret_cast<const float2 *>(p);
        v[0] = a.x; v[1] = a.y;
    }
}

template <typename TB, int VW, int L, int R, int UNR>
__global__ __launch_bounds__(kBlock) void k_row_blocked(RowBlockLaunch a) {
    constexpr int G = kWave / L;                 // row blocks per wavefront
    constexpr int WPB = kBlock / kWave;          // wavefronts per workgroup
    constexpr int WG = kRbWin / G;               // list entries per group and window
    static_assert(R == 2 || R == 4 || R == 8, "R");
    __shared__ int2 s_list[WPB][kRbWin];
    __shared__ __align__(16) float s_vals[WPB][kRbWin * R];
    const int t = threadIdx.x, wv = t / kWave, lane = t % kWave, g = lane / L, l = lane % L;
    const int blk_w = (xcd_remap(blockIdx.x, gridDim.x) * WPB + wv) * G;
    if (blk_w >= a.nblocks) return;              // the whole wavefront leaves
    const int blk = blk_w + g;
    const int col0 = blockIdx.y * a.slab_w + l * VW;
    const bool has = blk < a.nblocks;
    const bool active = has && col0 < a.nv;
    int u0 = 0, len = 0;
    if (has) {
        u0 = a.boff[blk];
        len = a.boff[blk + 1] - u0;
    }
    int wlen = len;                              // longest list of the wavefront
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) wlen = max(wlen, __shfl_xor(wlen, o));
    const TB *__restrict__ B = reinterpret_cast<const TB *>(a.B) + col0;
    int2 *sl = s_list[wv] + g * WG;
    float *sv = s_vals[wv] + g * WG * R;
    float acc[R][VW];
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int k = 0; k < VW; k++) acc[i][k] = 0.f;
    for (int w = 0; w < wlen; w += WG) {
        const int cnt = max(0, min(WG, len - w));
        wave_lds_sync();
        for (int i = l; i < cnt; i += L) sl[i] = a.list[u0 + w + i];
        if constexpr (R >= 4) {
            const float4 *src = reinterpret_cast<const float4 *>(a.vals + (int64_t)(u0 + w) * R);
            float4 *dst = reinterpret_cast<float4 *>(sv);
            for (int i = l; i < cnt * (R / 4); i += L) dst[i] = src[i];
        } else {
            const float2 *src = reinterpret_cast<const float2 *>(a.vals + (int64_t)(u0 + w) * R);
            float2 *dst = reinterpret_cast<float2 *>(sv);
            for (int i = l; i < cnt; i += L) dst[i] = src[i];
        }
        wave_lds_sync();
        if (!active) continue;
        for (int u = 0; u < cnt; u += UNR) {
            int2 cm[UNR];
            float bv[UNR][VW];
#pragma unroll
            for (int x = 0; x < UNR; x++) {
                cm[x] = sl[min(u + x, cnt - 1)];         // padded: no branch around loads
                if (u + x >= cnt) cm[x].y = 0;
            }
#pragma unroll
            for (int x = 0; x < UNR; x++) load_b<TB, VW>(bv[x], B + (int64_t)cm[x].x * a.ldb);
#pragma unroll
            for (int x = 0; x < UNR; x++) {
                float v[R];
                lds_row_vals<R>(v, sv + min(u + x, cnt - 1) * R);
#pragma unroll
                for (int i = 0; i < R; i++) {
                    // a branch, not a select: groups of one wavefront usually
                    // share the mask pattern, so rows without the column cost
                    // no vector instructions
                    if ((cm[x].y >> i) & 1) {
#pragma unroll
                        for (int k = 0; k < VW; k++) acc[i][k] = acc[i][k] + v[i] * bv[x][k];
                    }
                }
            }
        }
    }
    if (!active) return;
    const int64_t r0 = (int64_t)(blk / a.stride) * R * a.stride + blk % a.stride;
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int64_t r = r0 + (int64_t)i * a.stride;
        if (r < a.nr) store_vec_nt<VW>(a.C + r * a.ldc + col0, acc[i]);
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
        hipLaunchKernelGGL(k_split_combine, dim3(combine_blocks(a.nsplits, a.item.nv)), dim3(kBlock),
                           0, st, a.item, a.splits, a.nsplits);
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
                           dim3(kBlock), 0, st, a.items, a.split_prefix, a.count, a.nsplits);
        return launch_status();
    }
    return DSPMM_STATUS_SUCCESS;
}

namespace {
template <typename TB>
dspmm_status_t launch_spmv_e(const NbLaunch &a, hipStream_t st) {
    const dim3 g(a.ntiles), b(kBlock);
    switch (a.item.S) {
    case 512: hipLaunchKernelGGL((k_spmv_tiled<TB, 2>), g, b, 0, st, a.item, a.ntiles); break;
    case 1024: hipLaunchKernelGGL((k_spmv_tiled<TB, 4>), g, b, 0, st, a.item, a.ntiles); break;
    case 2048: hipLaunchKernelGGL((k_spmv_tiled<TB, 8>), g, b, 0, st, a.item, a.ntiles); break;
    case 4096: hipLaunchKernelGGL((k_spmv_tiled<TB, 16>), g, b, 0, st, a.item, a.ntiles); break;
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
    return launch_status();
}
}  // namespace

size_t nnz_cached_lds_bytes(int umax, int smax, int nv, int vw, bool bf16) {
    const size_t ls = (size_t)cached_stride(nv, vw);
    const size_t eb = bf16 ? 2 : 4;
    const size_t brows = (size_t)umax * ls * eb;
    // direct tiles: G = NT / L partial rows of ls floats, ls <= (L + 1) * vw, NT <= 256
    const size_t direct = (size_t)2 * kBlock * (size_t)vw * 4;
    const size_t stage = (size_t)smax * 4 + (size_t)((smax + 1) & ~1) * 2 + (size_t)kCachedRowsMax * 4;
    return std::max(brows, direct) + stage;
}

namespace {
template <typename TB, int VW, int L, int NT>
dspmm_status_t launch_cached_vln(const CachedLaunch &a, hipStream_t st) {
    if constexpr (L > NT) {
        return DSPMM_STATUS_NOT_SUPPORTED;
    } else {
        const size_t lds = nnz_cached_lds_bytes(a.umax, a.smax, a.nv, VW, sizeof(TB) == 2);
        hipLaunchKernelGGL((k_nnz_cached<TB, VW, L, NT>), dim3(a.ntiles), dim3(NT), lds, st, a);
        return launch_status();
    }
}
template <typename TB, int VW, int L>
dspmm_status_t launch_cached_vl(const CachedLaunch &a, hipStream_t st) {
    switch (a.threads) {
    case 64: return launch_cached_vln<TB, VW, L, 64>(a, st);
    case 128: return launch_cached_vln<TB, VW, L, 128>(a, st);
    case 256: return launch_cached_vln<TB, VW, L, 256>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}
template <typename TB, int VW>
dspmm_status_t launch_cached_l(const CachedLaunch &a, hipStream_t st) {
    switch (a.lanes) {
    case 1: return launch_cached_vl<TB, VW, 1>(a, st);
    case 2: return launch_cached_vl<TB, VW, 2>(a, st);
    case 4: return launch_cached_vl<TB, VW, 4>(a, st);
    case 8: return launch_cached_vl<TB, VW, 8>(a, st);
    case 16: return launch_cached_vl<TB, VW, 16>(a, st);
    case 32: return launch_cached_vl<TB, VW, 32>(a, st);
    case 64: return launch_cached_vl<TB, VW, 64>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}
}  // namespace

dspmm_status_t launch_nnz_cached(const CachedLaunch &a, hipStream_t st) {
    if (a.ntiles == 0) return DSPMM_STATUS_SUCCESS;
    if (a.lanes * a.vw < a.nv) return DSPMM_STATUS_NOT_SUPPORTED;
    if (a.bf16) {
        switch (a.vw) {
        case 8: return launch_cached_l<uint16_t, 8>(a, st);
        case 4: return launch_cached_l<uint16_t, 4>(a, st);
        case 2: return launch_cached_l<uint16_t, 2>(a, st);
        case 1: return launch_cached_l<uint16_t, 1>(a, st);
        default: return DSPMM_STATUS_NOT_SUPPORTED;
        }
    }
    switch (a.vw) {
    case 4: return launch_cached_l<float, 4>(a, st);
    case 2: return launch_cached_l<float, 2>(a, st);
    case 1: return launch_cached_l<float, 1>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}

namespace {
template <typename TB, int VW, int L, int R>
dspmm_status_t launch_rb_r(const RowBlockLaunch &a, hipStream_t st) {
    constexpr int G = kBlock / L;              // row blocks per workgroup
    const dim3 grid(div_up(a.nblocks, G), a.nslabs);
    if (a.unroll <= 4) hipLaunchKernelGGL((k_row_blocked<TB, VW, L, R, 4>), grid, dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((k_row_blocked<TB, VW, L, R, 8>), grid, dim3(kBlock), 0, st, a);
    return launch_status();
}
template <typename TB, int VW, int L>
dspmm_status_t launch_rb_l(const RowBlockLaunch &a, hipStream_t st) {
    switch (a.rows) {
    case 2: return launch_rb_r<TB, VW, L, 2>(a, st);
    case 4: return launch_rb_r<TB, VW, L, 4>(a, st);
    case 8: return launch_rb_r<TB, VW, L, 8>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}
template <typename TB, int VW>
dspmm_status_t launch_rb_v(const RowBlockLaunch &a, hipStream_t st) {
    switch (a.lanes) {
    case 1: return launch_rb_l<TB, VW, 1>(a, st);
    case 2: return launch_rb_l<TB, VW, 2>(a, st);
    case 4: return launch_rb_l<TB, VW, 4>(a, st);
    case 8: return launch_rb_l<TB, VW, 8>(a, st);
    case 16: return launch_rb_l<TB, VW, 16>(a, st);
    case 32: return launch_rb_l<TB, VW, 32>(a, st);
    case 64: return launch_rb_l<TB, VW, 64>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}
}  // namespace

dspmm_status_t launch_row_blocked(const RowBlockLaunch &a, hipStream_t st) {
    if (a.nblocks == 0) return DSPMM_STATUS_SUCCESS;
    if (a.bf16) {
        switch (a.vw) {
        case 8: return launch_rb_v<uint16_t, 8>(a, st);
        case 4: return launch_rb_v<uint16_t, 4>(a, st);
        case 2: return launch_rb_v<uint16_t, 2>(a, st);
        case 1: return launch_rb_v<uint16_t, 1>(a, st);
        default: return DSPMM_STATUS_NOT_SUPPORTED;
        }
    }
    switch (a.vw) {
    case 4: return launch_rb_v<float, 4>(a, st);
    case 2: return launch_rb_v<float, 2>(a, st);
    case 1: return launch_rb_v<float, 1>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}

// ===========================================================================
// Row-sequential SpMM (CSR_SCALAR with an unroll, and its BF16-B form).  A
// group of 2^lg lanes owns one row, lane j of the group VW consecutive
// columns of the slab; a wavefront owns G = 64 >> lg consecutive rows.  The
// wavefront stages its rows' (col, val) into its own LDS region: in one
// coalesced pass when they fit kRsWin pairs, else in windows split into one
// slice per group that still has pairs (so no group waits while another
// group's row is staged, and a long row gets the whole window once the
// others are done).  Each group then reads its row's pairs from
// LDS as broadcasts, issues UNR gathers of B and adds the products in
// storage order, so results equal the host loop's row sums.
// The last chunk of a window is padded with its last entry (loads without
// a branch keep the gathers back to back) and the padded products are
// skipped.  No workgroup barriers; every row is overwritten (empty rows
// -> 0).
// ===========================================================================
constexpr int kRsWin = kRowSeqWindow;

template <typename TB, int VW, int UNR, bool SLICED>
__global__ __launch_bounds__(kBlock) void k_row_seq(RowSeqLaunch g, int lg) {
    constexpr int WPB = kBlock / kWave;
    __shared__ int s_col[WPB][kRsWin];
    __shared__ float s_val[WPB][kRsWin];
    __shared__ int s_tab[WPB][SLICED ? 2 * kWave : 1];   // per window: slice start, count per rank
    const int t = threadIdx.x, wv = t / kWave, lane = t % kWave;
    const int G = kWave >> lg;
    const int64_t rw = ((int64_t)xcd_remap(blockIdx.x, gridDim.x) * WPB + wv) * G;
    if (rw >= g.nr) return;                      // the whole wavefront leaves
    const int nrow = (int)min((int64_t)G, (int64_t)g.nr - rw);
    const int gi = lane >> lg;
    const int64_t r = rw + gi;
    const int col0 = blockIdx.y * g.slab_w + (lane & ((1 << lg) - 1)) * VW;
    const bool active = gi < nrow && col0 < g.nv;
    const int e0 = g.rowPtr[rw], e1 = g.rowPtr[rw + nrow];
    int start = 0, end = 0;
    if (gi < nrow) {
        start = g.rowPtr[r];
        end = g.rowPtr[r + 1];
    }
    const TB *__restrict__ B = reinterpret_cast<const TB *>(g.B) + col0;
    float acc[VW];
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = 0.f;
    // Adds products p0 .. p1-1 of the staged pairs, UNR gathers in flight.
    auto run = [&](const int *sc, const float *sv, int p0, int p1) {
        for (int p = p0; p < p1; p += UNR) {
            float bv[UNR][VW];
#pragma unroll
            for (int u = 0; u < UNR; u++)
                load_b<TB, VW>(bv[u], B + (int64_t)sc[min(p + u, p1 - 1)] * g.ldb);
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                if (p + u < p1) {
                    const float a = sv[p + u];
#pragma unroll
                    for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * bv[u][k];
                }
            }
        }
    };
    if (!SLICED || e1 - e0 <= kRsWin) {
        // windows of the wavefront's contiguous pairs, one coalesced staging
        // pass each (a single window when they fit)
        int *sc = s_col[wv];
        float *sv = s_val[wv];
        for (int w = e0; w < e1; w += kRsWin) {
            const int cnt = min(kRsWin, e1 - w);
            wave_lds_sync();
            for (int i = lane; i < cnt; i += kWave) {
                sc[i] = ld_stream(g.colIdx + w + i);
                sv[i] = ld_stream(g.values + w + i);
            }
            wave_lds_sync();
            if (active) run(sc, sv, max(start, w) - w, min(end, w + cnt) - w);
        }
    } else if constexpr (SLICED) {
        // Windows shared by the groups that still have pairs: with n such
        // groups each gets a slice of sl = kRsWin / n (rounded down to a
        // power of two) pairs, so every group works in every window and a
        // long row gets the whole window once the short ones are done.  The
        // wavefront stages all slices together (lane i takes pair i of the
        // concatenated slices), coalesced within each slice.
        int *tab_s = s_tab[wv], *tab_c = s_tab[wv] + kWave;
        const bool leader = (lane & ((1 << lg) - 1)) == 0;
        const uint64_t below = (1ull << (gi << lg)) - 1;   // lanes before this group's leader
        const int len = end - start;
        int done = 0;
        for (;;) {
            const int rem = len - done;
            const uint64_t m = __ballot(leader && rem > 0);
            if (m == 0) break;
            const int nact = __popcll(m);
            int lsl = 9;                                    // log2 kRsWin
            while ((nact << lsl) > kRsWin) lsl--;
            const int sl = 1 << lsl;
            const int rank = __popcll(m & below);
            const int cnt = rem > 0 ? min(sl, rem) : 0;
            wave_lds_sync();
            if (leader && rem > 0) {
                tab_s[rank] = start + done;
                tab_c[rank] = cnt;
            }
            wave_lds_sync();
            for (int i = lane; i < (nact << lsl); i += kWave) {
                const int q = i >> lsl, o = i & (sl - 1);
                if (o < tab_c[q]) {
                    s_col[wv][i] = ld_stream(g.colIdx + tab_s[q] + o);
                    s_val[wv][i] = ld_stream(g.values + tab_s[q] + o);
                }
            }
            wave_lds_sync();
            if (active && cnt > 0) run(s_col[wv] + (rank << lsl), s_val[wv] + (rank << lsl), 0, cnt);
            done += cnt;
        }
    }
    if (active) store_vec_nt<VW>(g.C + r * g.ldc + col0, acc);
}

namespace {
template <typename TB, int VW, bool SL>
dspmm_status_t launch_row_seq_vs(const RowSeqLaunch &a, hipStream_t st) {
    int lg = 0;
    while ((1 << lg) < a.lanes) lg++;
    const dim3 grid(div_up(a.nr, kBlock >> lg), a.nslabs);
    switch (a.unroll) {
    case 2: hipLaunchKernelGGL((k_row_seq<TB, VW, 2, SL>), grid, dim3(kBlock), 0, st, a, lg); break;
    case 4: hipLaunchKernelGGL((k_row_seq<TB, VW, 4, SL>), grid, dim3(kBlock), 0, st, a, lg); break;
    case 8: hipLaunchKernelGGL((k_row_seq<TB, VW, 8, SL>), grid, dim3(kBlock), 0, st, a, lg); break;
    case 16: hipLaunchKernelGGL((k_row_seq<TB, VW, 16, SL>), grid, dim3(kBlock), 0, st, a, lg); break;
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
    return launch_status();
}
template <typename TB, int VW>
dspmm_status_t launch_row_seq_v(const RowSeqLaunch &a, hipStream_t st) {
    return a.sliced ? launch_row_seq_vs<TB, VW, true>(a, st) : launch_row_seq_vs<TB, VW, false>(a, st);
}
}  // namespace

bool row_seq_unroll_ok(int u) { return u == 2 || u == 4 || u == 8 || u == 16; }

dspmm_status_t launch_row_seq(const RowSeqLaunch &a, hipStream_t st) {
    if (a.nr == 0 || a.nv == 0) return DSPMM_STATUS_SUCCESS;
    if (!row_seq_unroll_ok(a.unroll) || a.lanes < 1 || a.lanes > 64 || (a.lanes & (a.lanes - 1)) ||
        a.slab_w != a.lanes * a.vw || (int64_t)a.nslabs * a.slab_w < a.nv)
        return DSPMM_STATUS_NOT_SUPPORTED;
    if (a.bf16) {
        switch (a.vw) {
        case 8: return launch_row_seq_v<uint16_t, 8>(a, st);
        case 4: return launch_row_seq_v<uint16_t, 4>(a, st);
        case 2: return launch_row_seq_v<uint16_t, 2>(a, st);
        case 1: return launch_row_seq_v<uint16_t, 1>(a, st);
        default: return DSPMM_STATUS_NOT_SUPPORTED;
        }
    }
    switch (a.vw) {
    case 4: return launch_row_seq_v<float, 4>(a, st);
    case 2: return launch_row_seq_v<float, 2>(a, st);
    case 1: return launch_row_seq_v<float, 1>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}

bool spmv_tile_ok(int T) { return T == 512 || T == 1024 || T == 2048 || T == 4096; }

dspmm_status_t launch_spmv_tiled(const NbLaunch &a, hipStream_t st) {
    if (a.ntiles == 0) return DSPMM_STATUS_SUCCESS;
    dspmm_status_t s = a.bf16 ? launch_spmv_e<uint16_t>(a, st) : launch_spmv_e<float>(a, st);
    if (s != DSPMM_STATUS_SUCCESS) return s;
    if (a.mid_event) (void)hipEventRecord(a.mid_event, st);
    if (a.nsplits > 0) {
        hipLaunchKernelGGL(k_split_combine, dim3(combine_blocks(a.nsplits, a.item.nv)), dim3(kBlock),
                           0, st, a.item, a.splits, a.nsplits);
        s = launch_status();
    }
    return s;
}

// dst[c * ldd + r] = src[r * lds + c] for r < rows, c < cols: 32 x 32 tiles
// through LDS (one pad column against bank conflicts), one tile per
// workgroup, tiles numbered row-major over the source.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_transpose(const T *__restrict__ src, int64_t lds, int rows,
                                                      int cols, T *__restrict__ dst, int64_t ldd) {
    __shared__ T tile[32][33];
    const int tiles_c = (cols + 31) / 32;
    const int tr = (int)(blockIdx.x / (unsigned)tiles_c), tc = (int)(blockIdx.x % (unsigned)tiles_c);
    const int r0 = tr * 32, c0 = tc * 32;
    const int tx = (int)threadIdx.x % 32, ty = (int)threadIdx.x / 32;
    for (int k = ty; k < 32; k += kBlock / 32) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[(int64_t)r * lds + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += kBlock / 32) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[(int64_t)c * ldd + r] = tile[tx][k];
    }
}

dspmm_status_t launch_transpose(const void *src, int64_t lds, int rows, int cols, void *dst,
                                int64_t ldd, int elem_bytes, hipStream_t st) {
    if (rows == 0 || cols == 0) return DSPMM_STATUS_SUCCESS;
    const int64_t tiles = (int64_t)((rows + 31) / 32) * ((cols + 31) / 32);
    if (tiles > INT_MAX) return DSPMM_STATUS_NOT_SUPPORTED;
    if (elem_bytes == 2)
        hipLaunchKernelGGL(k_transpose<uint16_t>, dim3((unsigned)tiles), dim3(kBlock), 0, st,
                           reinterpret_cast<const uint16_t *>(src), lds, rows, cols,
                           reinterpret_cast<uint16_t *>(dst), ldd);
    else
        hipLaunchKernelGGL(k_transpose<float>, dim3((unsigned)tiles), dim3(kBlock), 0, st,
                           reinterpret_cast<const float *>(src), lds, rows, cols,
                           reinterpret_cast<float *>(dst), ldd);
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
