// compat_kernels.hip -- the reference's 12-kernel SpMV/SpMM family,
// redesigned for gfx950 wave64, behind the reference's launcher names.
//
// Algorithm classes and output semantics follow the reference
// (spmvspmm.h:72-197): CSR = row-balanced, overwrite; COO = nnz-balanced,
// atomic accumulate.  What changes is the mapping onto CDNA4:
//   * lane groups of L lanes x VW floats cover a dense row slab (float4
//     along N) instead of one lane per column;
//   * the sparse operand of a workgroup is staged through LDS with
//     coalesced loads, then read as LDS broadcasts;
//   * parallel reductions are wave64 butterflies / segmented scans on
//     DPP, ds_swizzle and permlane32 (wave64.hpp);
//   * blocks are renumbered so each XCD walks a contiguous row range.
// Sequential (SR) kernels sum every output in nonzero order with a
// separately rounded multiply and add (the library is built with
// -ffp-contract=off), so the row-balanced SR kernels reproduce the
// reference host loop (spmvspmm_benchmark.cu:11-39) bit for bit.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "common.hpp"
#include "internal.hpp"

namespace dspmm {

constexpr int kWin = 2048;   // LDS staging window (nonzeros) of the row-balanced kernels

// Per-lane in-order accumulation of a window of nonzeros [p0, p1) staged
// in LDS at s_col/s_val (offset w), gathering VW floats of B per nonzero.
// Gathers are issued UNR at a time before they are summed, in order.
template <typename TB, int VW, int UNR>
__device__ __forceinline__ void seq_accumulate(float (&acc)[VW], int p0, int p1, int w,
                                               const int *s_col, const float *s_val,
                                               const TB *__restrict__ B, int64_t ldb,
                                               int col0) {
    int p = p0;
    for (; p + UNR <= p1; p += UNR) {
        float b[UNR][VW];
#pragma unroll
        for (int u = 0; u < UNR; u++)
            load_b<TB, VW>(b[u], B + (int64_t)s_col[p + u - w] * ldb + col0);
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const float a = s_val[p + u - w];
#pragma unroll
            for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * b[u][k];
        }
    }
    for (; p < p1; p++) {
        float b[VW];
        load_b<TB, VW>(b, B + (int64_t)s_col[p - w] * ldb + col0);
        const float a = s_val[p - w];
#pragma unroll
        for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * b[k];
    }
}

// Cooperative staging of nonzeros [w, w+cnt) into LDS by a whole block.
__device__ __forceinline__ void stage_window(int *s_col, float *s_val, const int *__restrict__ colIdx,
                                             const float *__restrict__ values, int w, int cnt) {
    for (int i = threadIdx.x; i < cnt; i += kBlock) {
        s_col[i] = ld_stream(colIdx + w + i);
        s_val[i] = ld_stream(values + w + i);
    }
}

// ===========================================================================
// Row-balanced, sequential reduction, row-major (spmv_csr_scalar for nv=1,
// spmm_csr_scalar_row for nv>1).  A group of L lanes owns one row; each lane
// owns VW consecutive columns of the slab.  The block's rows' nonzeros are
// staged through LDS in windows; every lane reads its row's (col, val) as an
// LDS broadcast within its group.  Overwrites every row (empty rows -> 0).
// SLICED: a block whose rows overflow one window, and one of whose rows is
// at least 4 shares of a window long (kWin / rows per block), splits each
// window into one slice per row that still has nonzeros (kWin / n for n
// such rows, rounded down to a power of two), staged by the whole block, so
// that no row waits while windows cover only other rows; a long row gets
// the whole window once the others are done.  Both forms add each row's
// products in nonzero order.
// ===========================================================================
template <int VW, int L, bool SLICED>
__global__ __launch_bounds__(kBlock) void k_csr_scalar_rm(int nr, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowPtr,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C, int slab_w) {
    constexpr int RPB = kBlock / L;
    constexpr int NW = kBlock / kWave;
    __shared__ int s_col[kWin];
    __shared__ float s_val[kWin];
    __shared__ int s_tab[SLICED ? 2 * RPB + NW : 1];   // slice start, count per rank; per-wave counts
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = b * RPB;
    const int r = r0 + (int)threadIdx.x / L;
    const int col0 = blockIdx.y * slab_w + ((int)threadIdx.x % L) * VW;
    const bool colok = col0 < nv;
    const int rlast = min(r0 + RPB, nr);
    const int bs = rowPtr[r0], be = rowPtr[rlast];
    int start = 0, end = 0;
    if (r < nr) { start = rowPtr[r]; end = rowPtr[r + 1]; }
    float acc[VW];
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = 0.f;
    bool slice = false;
    if constexpr (SLICED) {
        // slices pay where a row is long against its share of a window (kWin /
        // RPB pairs): then windows of contiguous pairs cover few rows each and
        // the others wait.  Rows shorter than 4 shares run in plain windows.
        if (be - bs > kWin) {
            int mx = end - start;
#pragma unroll
            for (int o = kWave / 2; o >= 1; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
            if ((int)threadIdx.x % kWave == 0) s_tab[2 * RPB + (int)threadIdx.x / kWave] = mx;
            __syncthreads();
            for (int k = 0; k < NW; k++) mx = max(mx, s_tab[2 * RPB + k]);
            slice = mx >= 4 * (kWin / RPB);
            __syncthreads();
        }
    }
    if (!slice) {
        for (int w = bs; w < be; w += kWin) {
            const int cnt = min(kWin, be - w);
            __syncthreads();
            stage_window(s_col, s_val, colIdx, values, w, cnt);
            __syncthreads();
            const int p0 = max(start, w), p1 = min(end, w + cnt);
            if (colok) seq_accumulate<float, VW, 4>(acc, p0, p1, w, s_col, s_val, B, ldb, col0);
        }
    } else if constexpr (SLICED) {
        const int lane = (int)threadIdx.x % kWave, wv = (int)threadIdx.x / kWave;
        const bool leader = (threadIdx.x % L) == 0;
        const uint64_t below = (1ull << (lane & ~(L - 1))) - 1;   // lanes before this group's leader
        int *tab_s = s_tab, *tab_c = s_tab + RPB, *wcnt = s_tab + 2 * RPB;
        const int len = end - start;
        int done = 0;
        for (;;) {
            const int rem = len - done;
            const uint64_t m = __ballot(leader && rem > 0);
            if (lane == 0) wcnt[wv] = __popcll(m);
            __syncthreads();
            int nact = 0, base = 0;
#pragma unroll
            for (int k = 0; k < NW; k++) {
                const int c = wcnt[k];
                nact += c;
                if (k < wv) base += c;
            }
            if (nact == 0) break;                         // the same in every thread
            int lsl = 0;
            while ((2 << lsl) * nact <= kWin) lsl++;      // largest power of two with nact * 2^lsl <= kWin
            const int rank = base + __popcll(m & below);
            const int cnt = rem > 0 ? min(1 << lsl, rem) : 0;
            if (leader && rem > 0) {
                tab_s[rank] = start + done;
                tab_c[rank] = cnt;
            }
            __syncthreads();
            for (int i = threadIdx.x; i < (nact << lsl); i += kBlock) {
                const int q = i >> lsl, o = i & ((1 << lsl) - 1);
                if (o < tab_c[q]) {
                    s_col[i] = ld_stream(colIdx + tab_s[q] + o);
                    s_val[i] = ld_stream(values + tab_s[q] + o);
                }
            }
            __syncthreads();
            if (colok && cnt > 0)
                seq_accumulate<float, VW, 4>(acc, rank << lsl, (rank << lsl) + cnt, 0, s_col, s_val,
                                             B, ldb, col0);
            done += cnt;
            __syncthreads();                              // before the next window's tables
        }
    }
    if (r < nr && colok) store_vec_nt<VW>(C + (int64_t)r * ldc + col0, acc);
}

// ===========================================================================
// Row-balanced, parallel reduction, row-major (spmv_csr_vector,
// spmm_csr_vector_row).  A team of GT groups x L lanes owns one row; group
// g takes nonzeros start+g, start+g+GT, ...; the GT partial vectors are
// combined with a butterfly over stride L..GT*L/2.  Overwrites.
// ===========================================================================
template <int VW, int L, int GT>
__global__ __launch_bounds__(kBlock) void k_csr_vector_rm(int nr, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowPtr,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C, int slab_w) {
    constexpr int T = GT * L;                  // lanes per row
    constexpr int RPB = kBlock / T;
    static_assert(T <= kWave, "team larger than a wavefront");
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int t = threadIdx.x;
    const int r = b * RPB + t / T;
    const int g = (t % T) / L;
    const int col0 = blockIdx.y * slab_w + (t % L) * VW;
    const bool colok = col0 < nv;
    float acc[VW];
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = 0.f;
    if (r < nr && colok) {
        const int s = rowPtr[r], e = rowPtr[r + 1];
        for (int p = s + g; p < e; p += GT) {
            const int c = colIdx[p];
            const float a = values[p];
            float bv[VW];
            load_vec<VW>(bv, B + (int64_t)c * ldb + col0);
#pragma unroll
            for (int k = 0; k < VW; k++) acc[k] = __builtin_fmaf(a, bv[k], acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < VW; k++) acc[k] = butterfly_sum<L, T>(acc[k]);
    if (r < nr && colok && g == 0) store_vec_nt<VW>(C + (int64_t)r * ldc + col0, acc);
}

// ===========================================================================
// Nnz-balanced, sequential reduction, row-major (spmv_coo_scalar,
// spmm_coo_scalar_row).  Each group of L lanes owns NE consecutive
// nonzeros; the block's nonzeros (row, col, val) are staged in LDS (padded
// so that groups reading their k-th element hit distinct banks).  A group
// sums runs of equal row in order and atomically adds each finished run.
// ===========================================================================
template <int VW, int L, int NE>
__global__ __launch_bounds__(kBlock) void k_coo_scalar_rm(int nnz, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowIdx,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C, int slab_w) {
    constexpr int GPB = kBlock / L;            // groups per block
    constexpr int CH = GPB * NE;               // nonzeros per block
    constexpr int PADN = CH + GPB;             // one pad word per group
    __shared__ int s_row[PADN];
    __shared__ int s_col[PADN];
    __shared__ float s_val[PADN];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t base = (int64_t)b * CH;
    const int cnt = (int)min<int64_t>(CH, (int64_t)nnz - base);
    for (int i = threadIdx.x; i < cnt; i += kBlock) {
        const int j = i + i / NE;
        s_row[j] = ld_stream(rowIdx + base + i);
        s_col[j] = ld_stream(colIdx + base + i);
        s_val[j] = ld_stream(values + base + i);
    }
    __syncthreads();
    const int grp = threadIdx.x / L;
    const int col0 = blockIdx.y * slab_w + ((int)threadIdx.x % L) * VW;
    const int e0 = grp * NE, e1 = min(e0 + NE, cnt);
    if (e0 >= e1 || col0 >= nv) return;
    const int j0 = grp * (NE + 1);
    int cur = s_row[j0];
    float acc[VW];
    {
        float bv[VW];
        load_vec<VW>(bv, B + (int64_t)s_col[j0] * ldb + col0);
        const float a = s_val[j0];
#pragma unroll
        for (int k = 0; k < VW; k++) acc[k] = a * bv[k];
    }
    constexpr int UNR = 4;
    int e = e0 + 1;
    for (; e < e1; e += UNR) {
        float bv[UNR][VW];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int j = j0 + (e + u - e0);
            if (e + u < e1) load_vec<VW>(bv[u], B + (int64_t)s_col[j] * ldb + col0);
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            if (e + u < e1) {
                const int j = j0 + (e + u - e0);
                const int rr = s_row[j];
                const float a = s_val[j];
                if (rr != cur) {
                    atomic_add_vec<VW>(C + (int64_t)cur * ldc + col0, acc);
#pragma unroll
                    for (int k = 0; k < VW; k++) acc[k] = a * bv[u][k];
                    cur = rr;
                } else {
#pragma unroll
                    for (int k = 0; k < VW; k++) acc[k] = acc[k] + a * bv[u][k];
                }
            }
        }
    }
    atomic_add_vec<VW>(C + (int64_t)cur * ldc + col0, acc);
}

// ===========================================================================
// Nnz-balanced, parallel reduction, row-major (spmv_coo_vector,
// spmm_coo_vector_row).  A wavefront takes G = 64/L consecutive nonzeros
// per step, one per group; products are combined by row with a wave64
// segmented scan (or a plain butterfly when the whole step is one row) and
// the last group of each row segment adds it atomically.  Grid-strided.
// ===========================================================================
template <int VW, int L>
__global__ __launch_bounds__(kBlock) void k_coo_vector_rm(int nnz, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowIdx,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C, int slab_w) {
    constexpr int G = kWave / L;
    const int ln = lane_id();
    const int g = ln / L;
    const int wave = (xcd_remap(blockIdx.x, gridDim.x) * kBlock + (int)threadIdx.x) / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
    const int col0 = blockIdx.y * slab_w + (ln % L) * VW;
    const bool colok = col0 < nv;
    for (int64_t base = (int64_t)wave * G; base < nnz; base += nwaves * G) {
        const int64_t p = base + g;
        const bool valid = p < nnz;
        int key = INT_MAX;
        float v[VW];
#pragma unroll
        for (int k = 0; k < VW; k++) v[k] = 0.f;
        if (valid) {
            key = rowIdx[p];
            if (colok) {
                const float a = values[p];
                float bv[VW];
                load_vec<VW>(bv, B + (int64_t)colIdx[p] * ldb + col0);
#pragma unroll
                for (int k = 0; k < VW; k++) v[k] = a * bv[k];
            }
        }
        const int kfirst = read_lane(key, 0), klast = read_lane(key, kWave - 1);
        if (kfirst == klast) {
#pragma unroll
            for (int k = 0; k < VW; k++) v[k] = butterfly_sum<L, kWave>(v[k]);
            if (g == 0 && colok) atomic_add_vec<VW>(C + (int64_t)key * ldc + col0, v);
        } else {
            segmented_scan<L, VW>(v, key);
            const bool tail = valid && (g == G - 1 || p + 1 >= nnz || rowIdx[p + 1] != key);
            if (tail && colok) atomic_add_vec<VW>(C + (int64_t)key * ldc + col0, v);
        }
    }
}

// ===========================================================================
// Column-major variants.  B(:,j) is contiguous (ld >= nc), C(:,j) too
// (ld >= nr).  The reference re-reads A once per column (grid.y = nv);
// here a lane carries T columns so A is read ceil(nv/T) times.
// ===========================================================================

// Row-balanced, sequential (spmm_csr_scalar_col): one lane per row.
template <int T>
__global__ __launch_bounds__(kBlock) void k_csr_scalar_cm(int nr, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowPtr,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C) {
    __shared__ int s_col[kWin];
    __shared__ float s_val[kWin];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = b * kBlock;
    const int r = r0 + (int)threadIdx.x;
    const int j0 = blockIdx.y * T;
    const int nt = min(T, nv - j0);
    const int rlast = min(r0 + kBlock, nr);
    const int bs = rowPtr[r0], be = rowPtr[rlast];
    int start = 0, end = 0;
    if (r < nr) { start = rowPtr[r]; end = rowPtr[r + 1]; }
    float acc[T];
#pragma unroll
    for (int t = 0; t < T; t++) acc[t] = 0.f;
    const float *Bj = B + (int64_t)j0 * ldb;
    for (int w = bs; w < be; w += kWin) {
        const int cnt = min(kWin, be - w);
        __syncthreads();
        stage_window(s_col, s_val, colIdx, values, w, cnt);
        __syncthreads();
        const int p0 = max(start, w), p1 = min(end, w + cnt);
        for (int p = p0; p < p1; p++) {
            const int c = s_col[p - w];
            const float a = s_val[p - w];
            float bv[T];
#pragma unroll
            for (int t = 0; t < T; t++) bv[t] = (t < nt) ? Bj[(int64_t)t * ldb + c] : 0.f;
#pragma unroll
            for (int t = 0; t < T; t++) acc[t] = acc[t] + a * bv[t];
        }
    }
    if (r < nr) {
#pragma unroll
        for (int t = 0; t < T; t++)
            if (t < nt) C[(int64_t)(j0 + t) * ldc + r] = acc[t];
    }
}

// Row-balanced, parallel (spmm_csr_vector_col, and spmv_csr_vector for
// column-major nv=1): L consecutive lanes per row stride its nonzeros.
template <int T, int L>
__global__ __launch_bounds__(kBlock) void k_csr_vector_cm(int nr, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowPtr,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C) {
    constexpr int RPB = kBlock / L;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int r = b * RPB + (int)threadIdx.x / L;
    const int l = (int)threadIdx.x % L;
    const int j0 = blockIdx.y * T;
    const int nt = min(T, nv - j0);
    const float *Bj = B + (int64_t)j0 * ldb;
    float acc[T];
#pragma unroll
    for (int t = 0; t < T; t++) acc[t] = 0.f;
    if (r < nr) {
        const int s = rowPtr[r], e = rowPtr[r + 1];
        for (int p = s + l; p < e; p += L) {
            const int c = colIdx[p];
            const float a = values[p];
#pragma unroll
            for (int t = 0; t < T; t++)
                if (t < nt) acc[t] = __builtin_fmaf(a, Bj[(int64_t)t * ldb + c], acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < T; t++) acc[t] = group_sum<L>(acc[t]);
    if (r < nr) {
        // Spread the T stores over the group's lanes.
#pragma unroll
        for (int t = 0; t < T; t++)
            if (t < nt && (t % L) == l) C[(int64_t)(j0 + t) * ldc + r] = acc[t];
    }
}

// Nnz-balanced, sequential (spmm_coo_scalar_col): one lane per NE nonzeros.
template <int T, int NE>
__global__ __launch_bounds__(kBlock) void k_coo_scalar_cm(int nnz, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowIdx,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C) {
    constexpr int CH = kBlock * NE;
    constexpr int PADN = CH + kBlock;
    __shared__ int s_row[PADN];
    __shared__ int s_col[PADN];
    __shared__ float s_val[PADN];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t base = (int64_t)b * CH;
    const int cnt = (int)min<int64_t>(CH, (int64_t)nnz - base);
    for (int i = threadIdx.x; i < cnt; i += kBlock) {
        const int j = i + i / NE;
        s_row[j] = ld_stream(rowIdx + base + i);
        s_col[j] = ld_stream(colIdx + base + i);
        s_val[j] = ld_stream(values + base + i);
    }
    __syncthreads();
    const int j0c = blockIdx.y * T;
    const int nt = min(T, nv - j0c);
    const float *Bj = B + (int64_t)j0c * ldb;
    float *Cj = C + (int64_t)j0c * ldc;
    const int e0 = (int)threadIdx.x * NE, e1 = min(e0 + NE, cnt);
    if (e0 >= e1) return;
    const int jb = (int)threadIdx.x * (NE + 1);
    int cur = s_row[jb];
    float acc[T];
    {
        const int c = s_col[jb];
        const float a = s_val[jb];
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = (t < nt) ? a * Bj[(int64_t)t * ldb + c] : 0.f;
    }
    for (int e = e0 + 1; e < e1; e++) {
        const int j = jb + (e - e0);
        const int rr = s_row[j], c = s_col[j];
        const float a = s_val[j];
        float bv[T];
#pragma unroll
        for (int t = 0; t < T; t++) bv[t] = (t < nt) ? Bj[(int64_t)t * ldb + c] : 0.f;
        if (rr != cur) {
#pragma unroll
            for (int t = 0; t < T; t++)
                if (t < nt) unsafeAtomicAdd(Cj + (int64_t)t * ldc + cur, acc[t]);
#pragma unroll
            for (int t = 0; t < T; t++) acc[t] = a * bv[t];
            cur = rr;
        } else {
#pragma unroll
            for (int t = 0; t < T; t++) acc[t] = acc[t] + a * bv[t];
        }
    }
#pragma unroll
    for (int t = 0; t < T; t++)
        if (t < nt) unsafeAtomicAdd(Cj + (int64_t)t * ldc + cur, acc[t]);
}

// Nnz-balanced, parallel (spmm_coo_vector_col, and spmv_coo_vector for
// column-major nv=1): a wavefront takes 64 consecutive nonzeros per step,
// T columns per lane, segmented scan per column.
template <int T>
__global__ __launch_bounds__(kBlock) void k_coo_vector_cm(int nnz, int nv, int64_t ldb, int64_t ldc,
                                                           const int *__restrict__ rowIdx,
                                                           const int *__restrict__ colIdx,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ B,
                                                           float *__restrict__ C) {
    const int ln = lane_id();
    const int wave = (xcd_remap(blockIdx.x, gridDim.x) * kBlock + (int)threadIdx.x) / kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (kBlock / kWave);
    const int j0 = blockIdx.y * T;
    const int nt = min(T, nv - j0);
    const float *Bj = B + (int64_t)j0 * ldb;
    float *Cj = C + (int64_t)j0 * ldc;
    for (int64_t base = (int64_t)wave * kWave; base < nnz; base += nwaves * kWave) {
        const int64_t p = base + ln;
        const bool valid = p < nnz;
        int key = INT_MAX;
        float v[T];
#pragma unroll
        for (int t = 0; t < T; t++) v[t] = 0.f;
        if (valid) {
            key = rowIdx[p];
            const int c = colIdx[p];
            const float a = values[p];
#pragma unroll
            for (int t = 0; t < T; t++)
                if (t < nt) v[t] = a * Bj[(int64_t)t * ldb + c];
        }
        const int kfirst = read_lane(key, 0), klast = read_lane(key, kWave - 1);
        if (kfirst == klast) {
#pragma unroll
            for (int t = 0; t < T; t++) v[t] = group_sum<kWave>(v[t]);
            // Lane t adds column t: one atomic per column, spread over lanes.
#pragma unroll
            for (int t = 0; t < T; t++)
                if (ln == t && t < nt) unsafeAtomicAdd(Cj + (int64_t)t * ldc + key, v[t]);
        } else {
            segmented_scan<1, T>(v, key);
            const bool tail = valid && (ln == kWave - 1 || p + 1 >= nnz || rowIdx[p + 1] != key);
            if (tail) {
#pragma unroll
                for (int t = 0; t < T; t++)
                    if (t < nt) unsafeAtomicAdd(Cj + (int64_t)t * ldc + key, v[t]);
            }
        }
    }
}

// ===========================================================================
// Host side: template selection and launch.
// ===========================================================================
namespace {

constexpr int kCmCols = 8;      // columns per lane in the column-major kernels
constexpr int kCooNE = 8;       // nonzeros per group in the COO scalar kernels

int grid_for_waves(int64_t waves_needed) {
    // Persistent-ish grid for the grid-strided COO vector kernels: enough
    // workgroups to fill MI355X many times over, capped by the work.
    const int64_t blocks = (waves_needed + 3) / 4;
    const int64_t cap = 256 * 64;
    return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, cap));
}

// Lanes per row for the parallel-reduction row-balanced kernels, from the
// mean row length (same thresholds as classic CSR-vector sub-wave sizing).
int lanes_for_mean(double mean, int cap) {
    int g = 2;
    if (mean > 4) g = 4;
    if (mean > 8) g = 8;
    if (mean > 16) g = 16;
    if (mean > 32) g = 32;
    if (mean > 64) g = 64;
    return std::min(g, cap);
}

template <typename F>
void with_lanes(int lanes, F &&f) {
    switch (lanes) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: f(std::integral_constant<int, 64>{}); break;
    }
}
template <typename F>
void with_vw(int vw, F &&f) {
    switch (vw) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    default: f(std::integral_constant<int, 1>{}); break;
    }
}

dspmm_status_t csr_scalar_rm(int nr, int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowPtr,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st) {
    const Slab s = make_slab(nv, choose_vw(nv, ldb, ldc, B, C, 4, 4));
    // sliced windows where a block's rows typically fill more than half a
    // window (the mean row length decides; a block that fits one window
    // runs the same code in both forms)
    const bool sliced = (double)nnz / std::max(1, nr) * (kBlock / s.lanes) > kWin / 2;
    with_vw(s.vw, [&](auto vw_c) {
        with_lanes(s.lanes, [&](auto l_c) {
            constexpr int VW_ = decltype(vw_c)::value, L_ = decltype(l_c)::value;
            dim3 grid(div_up(nr, kBlock / L_), s.count);
            if (sliced)
                hipLaunchKernelGGL((k_csr_scalar_rm<VW_, L_, true>), grid, dim3(kBlock), 0, st, nr,
                                   nv, ldb, ldc, rowPtr, colIdx, values, B, C, s.width);
            else
                hipLaunchKernelGGL((k_csr_scalar_rm<VW_, L_, false>), grid, dim3(kBlock), 0, st, nr,
                                   nv, ldb, ldc, rowPtr, colIdx, values, B, C, s.width);
        });
    });
    return launch_status();
}

template <int VW, int L, int GT>
void launch_csr_vector_rm(int nr, int nv, int64_t ldb, int64_t ldc, const int *rowPtr,
                          const int *colIdx, const float *values, const float *B, float *C,
                          hipStream_t st, const Slab &s) {
    if constexpr (GT * L <= kWave) {
        dim3 grid(div_up(nr, kBlock / (GT * L)), s.count);
        hipLaunchKernelGGL((k_csr_vector_rm<VW, L, GT>), grid, dim3(kBlock), 0, st, nr, nv, ldb,
                           ldc, rowPtr, colIdx, values, B, C, s.width);
    }
}

dspmm_status_t csr_vector_rm(int nr, int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowPtr,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st, int hint) {
    const Slab s = make_slab(nv, choose_vw(nv, ldb, ldc, B, C, 4, 4));
    const double mean = nr ? (double)nnz / nr : 0.0;
    int gt = hint > 0 ? pow2_floor(hint) : lanes_for_mean(mean, 64);
    gt = std::max(1, std::min(gt, 64 / s.lanes));
    with_vw(s.vw, [&](auto vw_c) {
        with_lanes(s.lanes, [&](auto l_c) {
            constexpr int VW_ = decltype(vw_c)::value, L_ = decltype(l_c)::value;
            with_lanes(gt, [&](auto gt_c) {
                constexpr int GT_ = decltype(gt_c)::value;
                launch_csr_vector_rm<VW_, L_, GT_>(nr, nv, ldb, ldc, rowPtr, colIdx, values, B, C,
                                                   st, s);
            });
        });
    });
    return launch_status();
}

dspmm_status_t coo_scalar_rm(int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowIdx,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st) {
    const Slab s = make_slab(nv, choose_vw(nv, ldb, ldc, B, C, 4, 4));
    with_vw(s.vw, [&](auto vw_c) {
        with_lanes(s.lanes, [&](auto l_c) {
            constexpr int VW_ = decltype(vw_c)::value, L_ = decltype(l_c)::value;
            constexpr int CH = (kBlock / L_) * kCooNE;
            dim3 grid(div_up(nnz, CH), s.count);
            hipLaunchKernelGGL((k_coo_scalar_rm<VW_, L_, kCooNE>), grid, dim3(kBlock), 0, st, nnz,
                               nv, ldb, ldc, rowIdx, colIdx, values, B, C, s.width);
        });
    });
    return launch_status();
}

dspmm_status_t coo_vector_rm(int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowIdx,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st) {
    const Slab s = make_slab(nv, choose_vw(nv, ldb, ldc, B, C, 4, 4));
    with_vw(s.vw, [&](auto vw_c) {
        with_lanes(s.lanes, [&](auto l_c) {
            constexpr int VW_ = decltype(vw_c)::value, L_ = decltype(l_c)::value;
            constexpr int G = kWave / L_;
            dim3 grid(grid_for_waves(((int64_t)nnz + G - 1) / G), s.count);
            hipLaunchKernelGGL((k_coo_vector_rm<VW_, L_>), grid, dim3(kBlock), 0, st, nnz, nv, ldb,
                               ldc, rowIdx, colIdx, values, B, C, s.width);
        });
    });
    return launch_status();
}

dspmm_status_t csr_scalar_cm(int nr, int nv, int64_t ldb, int64_t ldc, const int *rowPtr,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st) {
    dim3 grid(div_up(nr, kBlock), div_up(nv, kCmCols));
    hipLaunchKernelGGL((k_csr_scalar_cm<kCmCols>), grid, dim3(kBlock), 0, st, nr, nv, ldb, ldc,
                       rowPtr, colIdx, values, B, C);
    return launch_status();
}

dspmm_status_t csr_vector_cm(int nr, int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowPtr,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st, int hint) {
    const double mean = nr ? (double)nnz / nr : 0.0;
    const int L = hint > 0 ? std::min(64, pow2_floor(hint)) : lanes_for_mean(mean, 64);
    const int T = nv == 1 ? 1 : kCmCols;
    with_lanes(L, [&](auto l_c) {
        constexpr int L_ = decltype(l_c)::value;
        if (T == 1) {
            dim3 grid(div_up(nr, kBlock / L_), nv);
            hipLaunchKernelGGL((k_csr_vector_cm<1, L_>), grid, dim3(kBlock), 0, st, nr, nv, ldb,
                               ldc, rowPtr, colIdx, values, B, C);
        } else {
            dim3 grid(div_up(nr, kBlock / L_), div_up(nv, kCmCols));
            hipLaunchKernelGGL((k_csr_vector_cm<kCmCols, L_>), grid, dim3(kBlock), 0, st, nr, nv,
                               ldb, ldc, rowPtr, colIdx, values, B, C);
        }
    });
    return launch_status();
}

dspmm_status_t coo_scalar_cm(int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowIdx,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st) {
    dim3 grid(div_up(nnz, kBlock * kCooNE), div_up(nv, kCmCols));
    hipLaunchKernelGGL((k_coo_scalar_cm<kCmCols, kCooNE>), grid, dim3(kBlock), 0, st, nnz, nv, ldb,
                       ldc, rowIdx, colIdx, values, B, C);
    return launch_status();
}

dspmm_status_t coo_vector_cm(int nnz, int nv, int64_t ldb, int64_t ldc, const int *rowIdx,
                             const int *colIdx, const float *values, const float *B, float *C,
                             hipStream_t st) {
    if (nv == 1) {
        dim3 grid(grid_for_waves(((int64_t)nnz + kWave - 1) / kWave), 1);
        hipLaunchKernelGGL((k_coo_vector_cm<1>), grid, dim3(kBlock), 0, st, nnz, nv, ldb, ldc,
                           rowIdx, colIdx, values, B, C);
    } else {
        dim3 grid(grid_for_waves(((int64_t)nnz + kWave - 1) / kWave), div_up(nv, kCmCols));
        hipLaunchKernelGGL((k_coo_vector_cm<kCmCols>), grid, dim3(kBlock), 0, st, nnz, nv, ldb,
                           ldc, rowIdx, colIdx, values, B, C);
    }
    return launch_status();
}

}  // namespace

// Shared entry for the 12 launchers, the dispatcher and the native API.
dspmm_status_t launch_compat(int which, int nr, int nc, int nnz, int nv, int64_t ldb, int64_t ldc,
                             const int *rowPtr, const int *rowIdx, const int *colIdx,
                             const float *values, const float *B, float *C, hipStream_t st,
                             int hint) {
    (void)nc;
    const bool coo = which == K_SPMV_COO_SCALAR || which == K_SPMV_COO_VECTOR ||
                     which == K_SPMM_COO_SCALAR_ROW || which == K_SPMM_COO_VECTOR_ROW ||
                     which == K_SPMM_COO_SCALAR_COL || which == K_SPMM_COO_VECTOR_COL;
    if (nr == 0 || nv == 0) return DSPMM_STATUS_SUCCESS;
    if (coo && nnz == 0) return DSPMM_STATUS_SUCCESS;   // nothing to accumulate
    switch (which) {
    case K_SPMV_CSR_SCALAR:
    case K_SPMM_CSR_SCALAR_ROW:
        return csr_scalar_rm(nr, nnz, nv, ldb, ldc, rowPtr, colIdx, values, B, C, st);
    case K_SPMV_CSR_VECTOR:
    case K_SPMM_CSR_VECTOR_ROW:
        return csr_vector_rm(nr, nnz, nv, ldb, ldc, rowPtr, colIdx, values, B, C, st, hint);
    case K_SPMV_COO_SCALAR:
    case K_SPMM_COO_SCALAR_ROW:
        return coo_scalar_rm(nnz, nv, ldb, ldc, rowIdx, colIdx, values, B, C, st);
    case K_SPMV_COO_VECTOR:
    case K_SPMM_COO_VECTOR_ROW:
        return coo_vector_rm(nnz, nv, ldb, ldc, rowIdx, colIdx, values, B, C, st);
    case K_SPMM_CSR_SCALAR_COL:
        return csr_scalar_cm(nr, nv, ldb, ldc, rowPtr, colIdx, values, B, C, st);
    case K_SPMM_CSR_VECTOR_COL:
        return csr_vector_cm(nr, nnz, nv, ldb, ldc, rowPtr, colIdx, values, B, C, st, hint);
    case K_SPMM_COO_SCALAR_COL:
        return coo_scalar_cm(nnz, nv, ldb, ldc, rowIdx, colIdx, values, B, C, st);
    case K_SPMM_COO_VECTOR_COL:
        return coo_vector_cm(nnz, nv, ldb, ldc, rowIdx, colIdx, values, B, C, st);
    }
    return DSPMM_STATUS_INVALID_VALUE;
}

}  // namespace dspmm
