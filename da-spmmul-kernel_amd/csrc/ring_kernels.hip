// ring_kernels.hip -- k_ring: row-major SpMM through a ring of B rows held
// in LDS and filled by a host-computed schedule (ring_plan.cpp).
//
// One workgroup walks one stream of tiles.  The plan says, per tile, which
// B rows to load into which ring slots before the tile runs, and gives
// every nonzero the slot of its B row.  While the workgroup computes tile
// t from LDS, the 16-byte pieces of tile t+1's fills and its (slot, value)
// metadata are already in flight into registers; after a barrier they are
// written to LDS (the plan never evicts a row tile t+1 needs, and tile t is
// done with every slot by then).  A group of L lanes x VW columns sums one
// row at a time in its nonzero order, with a separately rounded multiply
// and add, so every row equals the reference host loop
// (spmvspmm_benchmark.cu:11-39) bit for bit, with FP32 or BF16-stored B.
//
// What this buys on a grid stencil: the row-gathering kernels fetch a B row
// through L1 for every nonzero (27 per output row on a 27-point stencil);
// here a workgroup's B traffic is its tiles' halos, about two B rows per
// output row, and the 27 reads per output row are LDS reads.
#include <algorithm>

#include "common.hpp"
#include "ring.hpp"

namespace dspmm {

constexpr int kRingFillRegs = 4;     // 16-byte fill pieces per thread in flight during a tile

__host__ __device__ inline size_t ring_bytes16(size_t b) { return (b + 15) / 16 * 16; }

// LDS layout: ring (slots + the zero slot) | ring byte offsets (int) and
// values (float) of the tile's nonzeros, 8 spare entries each | rows
// (rows_max int2) | stream tiles (int4) | stream fill offsets (int) |
// stream fills (int2).
struct RingLayout {
    size_t offs, vals, rows, tiles, foff, fills, total;
};
__host__ __device__ inline RingLayout ring_layout(int slots, int nv, bool bf16, int rows_max,
                                                  int nnz_max, int stiles, int sfills) {
    RingLayout l;
    l.offs = ring_bytes16((size_t)(slots + 1) * ring_row_elems(nv, bf16) * (bf16 ? 2 : 4));
    l.vals = l.offs + 4 * ((size_t)nnz_max + 8);
    l.rows = l.vals + 4 * ((size_t)nnz_max + 8);
    l.tiles = ring_bytes16(l.rows + 8 * (size_t)rows_max);
    l.foff = l.tiles + 16 * (size_t)stiles;
    l.fills = ring_bytes16(l.foff + 4 * ((size_t)stiles + 1));
    l.total = l.fills + 8 * ((size_t)sfills + 1);
    return l;
}

size_t ring_lds_bytes(const RingLaunch &a) {
    return ring_layout(a.slots, a.nv, a.bf16, a.rows_max, a.nnz_max, a.stream_tiles_max,
                       a.stream_fills_max).total;
}

namespace {

struct RingSmem {
    unsigned char *ring;     // (slots + 1) x ls elements; the last slot is zeros
    int *offs;               // per nonzero of the current tile: byte offset of its B row
    float *vals;             // per nonzero of the current tile: its value
    int2 *rows;              // {row, end offset} per row of the current tile
    int4 *tiles;             // the stream's tile descriptors
    int *foff;               // the stream's fill offsets (relative to its first)
    int2 *fills;             // the stream's fills {column, slot}
};

// A B row is `parts` 16-byte pieces (a power of two dividing the
// workgroup): thread t always loads piece q = t % parts, of fills
// t / parts, t / parts + NT / parts, ...  Fills [base, base + the chunk) of
// a tile whose fills are entries [fb, fb + nf) of the stream's list.  The
// loads are issued with no branch around them (a fill past the end reloads
// entry fb, which exists, and is marked dst = -1), so nothing waits for
// them until fill_store.
typedef unsigned int piece_t __attribute__((ext_vector_type(4)));
struct FillRegs {
    piece_t v[kRingFillRegs];
    int dst[kRingFillRegs];  // byte offset in the ring, -1: no piece
};

struct PieceMap {
    int q;                   // this thread's piece of every row it loads
    int f;                   // its first fill
    int step;                // fills per pass of the workgroup
};

template <typename TB>
__device__ __forceinline__ void fill_load(const RingLaunch &a, const RingSmem &s,
                                          const PieceMap &pm, int fb, int nf, int base, int ls,
                                          FillRegs &f) {
    constexpr int per = 16 / (int)sizeof(TB);         // elements per piece
    const TB *__restrict__ B = reinterpret_cast<const TB *>(a.B) + pm.q * per;
    int2 fl[kRingFillRegs];
#pragma unroll
    for (int i = 0; i < kRingFillRegs; i++) {
        const int k = base + pm.f + i * pm.step;
        fl[i] = s.fills[fb + (k < nf ? k : 0)];
        f.dst[i] = k < nf ? (fl[i].y * ls + pm.q * per) * (int)sizeof(TB) : -1;
    }
#pragma unroll
    for (int i = 0; i < kRingFillRegs; i++)
        f.v[i] = *reinterpret_cast<const piece_t *>(B + (int64_t)fl[i].x * a.ldb);
}

__device__ __forceinline__ void fill_store(const RingSmem &s, const FillRegs &f) {
#pragma unroll
    for (int i = 0; i < kRingFillRegs; i++)
        if (f.dst[i] >= 0) *reinterpret_cast<piece_t *>(s.ring + f.dst[i]) = f.v[i];
}

// Metadata of one tile: thread t holds nonzeros [8t, 8t + 8) of the tile
// (values and slots, 16-byte aligned in the plan) and row t.  Loaded with
// no branch: a thread past the end reloads the last chunk / row (the plan
// pads 8 entries after its last tile).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
struct MetaRegs {
    f32x4_t v0, v1;
    piece_t sl;
    int2 row;
    int nch, nrows;
};

__device__ __forceinline__ void meta_load(const RingLaunch &a, const int4 td, MetaRegs &m) {
    const int t = threadIdx.x;
    m.nch = (td.y - td.x) >> 3;
    m.nrows = td.w - td.z;
    const int c = min(t, max(m.nch - 1, 0));
    const f32x4_t *pv = reinterpret_cast<const f32x4_t *>(a.pval + td.x) + 2 * c;
    m.v0 = pv[0];
    m.v1 = pv[1];
    m.sl = reinterpret_cast<const piece_t *>(a.pslot + td.x)[c];
    m.row = a.row_info[td.z + min(t, m.nrows - 1)];
}

__device__ __forceinline__ void meta_store(const RingSmem &s, const MetaRegs &m, int row_bytes) {
    const int t = threadIdx.x;
    if (t < m.nch) {
        int4 *o = reinterpret_cast<int4 *>(s.offs + 8 * t);
        o[0] = make_int4((int)(m.sl[0] & 0xffffu) * row_bytes, (int)(m.sl[0] >> 16) * row_bytes,
                         (int)(m.sl[1] & 0xffffu) * row_bytes, (int)(m.sl[1] >> 16) * row_bytes);
        o[1] = make_int4((int)(m.sl[2] & 0xffffu) * row_bytes, (int)(m.sl[2] >> 16) * row_bytes,
                         (int)(m.sl[3] & 0xffffu) * row_bytes, (int)(m.sl[3] >> 16) * row_bytes);
        f32x4_t *v = reinterpret_cast<f32x4_t *>(s.vals + 8 * t);
        v[0] = m.v0;
        v[1] = m.v1;
    }
    if (t < m.nrows) s.rows[t] = m.row;
}

// A tile's fills from `base` on, synchronously (a stream's first tile,
// and fills beyond what the registers carried).
template <typename TB>
__device__ __forceinline__ void fill_sync(const RingLaunch &a, const RingSmem &s,
                                          const PieceMap &pm, int fb, int nf, int base, int ls) {
    for (int b = base; b < nf; b += pm.step * kRingFillRegs) {
        FillRegs f;
        fill_load<TB>(a, s, pm, fb, nf, b, ls, f);
        fill_store(s, f);
    }
}

// Group g sums rows g, g + G, ..., g + (R - 1) G of the tile side by side
// (a row past the tile's end recomputes its last row and is not stored),
// into registers; store_rows writes them later, after the next tile's LDS
// writes, so the stores' latency overlaps the next tile (on gfx9 stores
// count against the same counter as the prefetch loads).  All rows of a
// tile are padded to one length, a multiple of 4 (zero-slot entries), so
// the loop over a row's entries, four at a time (offsets and values in
// 16-byte reads), has no bounds inside.
template <int VW, int R>
struct TileRows {
    float acc[R][VW];
    int row[R];                       // -1: none
};

template <typename TB, int VW, int L, int R, int NT>
__device__ __forceinline__ void compute_tile(const RingLaunch &a, const RingSmem &s, int nrows,
                                             TileRows<VW, R> &out) {
    constexpr int G = NT / L;
    const int t = threadIdx.x, g = t / L;
    const int col0 = (t % L) * VW;
    // lanes past nv read a valid column of the row (nv is a multiple of
    // VW) and store nothing
    const unsigned char *rbase = s.ring + min(col0, a.nv - VW) * (int)sizeof(TB);
    const int len = s.rows[0].y;                    // every row's padded length
    int base[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int i = g + j * G;
        const int ie = min(i, nrows - 1);
        base[j] = ie * len;
        out.row[j] = i < nrows ? s.rows[ie].x : -1;
#pragma unroll
        for (int k = 0; k < VW; k++) out.acc[j][k] = 0.f;
    }
    for (int p = 0; p < len; p += 4) {
        int4 o[R];
        float4 v[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            o[j] = *reinterpret_cast<const int4 *>(s.offs + base[j] + p);
            v[j] = *reinterpret_cast<const float4 *>(s.vals + base[j] + p);
        }
        float b[R][4][VW];
#pragma unroll
        for (int j = 0; j < R; j++) {
            load_b<TB, VW>(b[j][0], reinterpret_cast<const TB *>(rbase + o[j].x));
            load_b<TB, VW>(b[j][1], reinterpret_cast<const TB *>(rbase + o[j].y));
            load_b<TB, VW>(b[j][2], reinterpret_cast<const TB *>(rbase + o[j].z));
            load_b<TB, VW>(b[j][3], reinterpret_cast<const TB *>(rbase + o[j].w));
        }
#pragma unroll
        for (int j = 0; j < R; j++)
#pragma unroll
            for (int k = 0; k < VW; k++) {
                float c = out.acc[j][k];
                c = c + v[j].x * b[j][0][k];
                c = c + v[j].y * b[j][1][k];
                c = c + v[j].z * b[j][2][k];
                c = c + v[j].w * b[j][3][k];
                out.acc[j][k] = c;
            }
    }
}

template <int VW, int L, int R>
__device__ __forceinline__ void store_rows(const RingLaunch &a, const TileRows<VW, R> &r) {
    const int col0 = (threadIdx.x % L) * VW;
    if (col0 >= a.nv) return;
#pragma unroll
    for (int j = 0; j < R; j++)
        if (r.row[j] >= 0) store_vec_nt<VW>(a.C + (int64_t)r.row[j] * a.ldc + col0, r.acc[j]);
}

}  // namespace

// Workgroups of NT = 256 or 512 threads.  The ring takes a large part of a
// CU's LDS, so few workgroups share a CU: registers are budgeted for two
// (NT / 128 waves per SIMD), keeping the prefetched fills in registers
// across the compute loop.
template <typename TB, int VW, int L, int R, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, NT / 128)))
void k_ring(RingLaunch a) {
    extern __shared__ __align__(16) unsigned char rsm[];
    const int ls = ring_row_elems(a.nv, sizeof(TB) == 2);
    const RingLayout lay = ring_layout(a.slots, a.nv, sizeof(TB) == 2, a.rows_max, a.nnz_max,
                                       a.stream_tiles_max, a.stream_fills_max);
    RingSmem s;
    s.ring = rsm;
    s.offs = reinterpret_cast<int *>(rsm + lay.offs);
    s.vals = reinterpret_cast<float *>(rsm + lay.vals);
    s.rows = reinterpret_cast<int2 *>(rsm + lay.rows);
    const int row_bytes = ls * (int)sizeof(TB);
    s.tiles = reinterpret_cast<int4 *>(rsm + lay.tiles);
    s.foff = reinterpret_cast<int *>(rsm + lay.foff);
    s.fills = reinterpret_cast<int2 *>(rsm + lay.fills);
    const int stream = xcd_remap(blockIdx.x, gridDim.x);
    const int t0 = a.stream_off[stream], t1 = a.stream_off[stream + 1];
    if (t0 >= t1) return;
    const int ns = t1 - t0;
    const int t = threadIdx.x;
    PieceMap pm;
    {
        const int parts = a.nv * (int)sizeof(TB) / 16;      // a power of two <= NT
        pm.q = t & (parts - 1);
        pm.f = t / parts;
        pm.step = NT / parts;
    }

    {   // the stream's tile descriptors and fill list, into LDS
        const int f0 = a.fill_off[t0], nf = a.fill_off[t1] - f0;
        for (int i = t; i < ns; i += NT) {
            s.tiles[i] = a.tiles[t0 + i];
            s.foff[i] = a.fill_off[t0 + i] - f0;
        }
        if (t == 0) s.foff[ns] = nf;
        for (int i = t; i <= nf; i += NT) s.fills[i] = a.fills[f0 + i];   // + the padding
        // the zero slot, which padding entries read
        for (int i = t; i < row_bytes / 4; i += NT)
            reinterpret_cast<int *>(s.ring + (size_t)a.slots * row_bytes)[i] = 0;
        __syncthreads();
    }
    int nrows;
    {   // the stream's first tile: fills and metadata, synchronously
        MetaRegs m;
        meta_load(a, s.tiles[0], m);
        fill_sync<TB>(a, s, pm, 0, s.foff[1], 0, ls);
        meta_store(s, m, row_bytes);
        nrows = m.nrows;
        __syncthreads();
    }
    for (int k = 0; k < ns; k++) {
        // tile k + 1's fills and metadata in flight while tile k runs (the
        // last tile reloads its own metadata, unused)
        const bool next = k + 1 < ns;
        const int kn = next ? k + 1 : k;
        const int fb = s.foff[kn];
        const int nf = next ? s.foff[kn + 1] - fb : 0;
        FillRegs f;
        MetaRegs m;
        fill_load<TB>(a, s, pm, fb, nf, 0, ls, f);
        meta_load(a, s.tiles[kn], m);
        TileRows<VW, R> out;
        compute_tile<TB, VW, L, R, NT>(a, s, nrows, out);
        if (!next) {
            store_rows<VW, L, R>(a, out);
            break;
        }
        __syncthreads();                  // tile k is done with its slots and metadata
        fill_store(s, f);
        meta_store(s, m, row_bytes);
        store_rows<VW, L, R>(a, out);
        fill_sync<TB>(a, s, pm, fb, nf, pm.step * kRingFillRegs, ls);
        nrows = m.nrows;
        __syncthreads();
    }
}

// Workgroup size: 512 threads where the LDS a workgroup needs leaves room
// for at most two per CU (more waves each), else 256 (more workgroups).
static int ring_threads(const RingLaunch &a) {
    if (ring_lds_bytes(a) * 3 > 160 * 1024 && (a.lanes == 8 || a.lanes == 16 || a.lanes == 32) &&
        (a.rows_max + 512 / a.lanes - 1) / (512 / a.lanes) <= 2 && a.nnz_max <= 8 * 512)
        return 512;
    return 256;
}

namespace {
template <typename TB, int VW, int L, int NT>
dspmm_status_t launch_ring_n(const RingLaunch &a, hipStream_t st) {
    constexpr int G = NT / L;
    const size_t lds = ring_lds_bytes(a);
    const int rows_per_group = (a.rows_max + G - 1) / G;
    if (rows_per_group <= 1)
        hipLaunchKernelGGL((k_ring<TB, VW, L, 1, NT>), dim3(a.nstreams), dim3(NT), lds, st, a);
    else if (rows_per_group == 2 || NT == 512)
        hipLaunchKernelGGL((k_ring<TB, VW, L, 2, NT>), dim3(a.nstreams), dim3(NT), lds, st, a);
    else
        hipLaunchKernelGGL((k_ring<TB, VW, L, (NT == 512 ? 2 : 4), NT>), dim3(a.nstreams),
                           dim3(NT), lds, st, a);
    return launch_status();
}
template <typename TB, int VW, int L>
dspmm_status_t launch_ring_vl(const RingLaunch &a, hipStream_t st) {
    if constexpr (L >= 8 && L <= 32)
        if (ring_threads(a) == 512) return launch_ring_n<TB, VW, L, 512>(a, st);
    return launch_ring_n<TB, VW, L, 256>(a, st);
}
}  // namespace

bool ring_supported(const RingLaunch &a) {
    const int eb = a.bf16 ? 2 : 4;
    const int nt = ring_threads(a);
    const int parts = a.nv * eb / 16;
    // one slab; every B row a whole number of 16-byte pieces at 16-byte
    // aligned addresses, a power of two of them; a tile's metadata in one
    // register load; its rows in the lane groups' registers
    if (a.vw != (a.bf16 ? 8 : 4) || a.lanes * a.vw < a.nv || (a.nv * eb) % 16 || a.nv % a.vw ||
        parts < 1 || parts > nt || (parts & (parts - 1)) || (a.ldb * eb) % 16 ||
        !aligned(a.B, 16) || !aligned(a.C, 16) || (a.ldc % a.vw) || a.nnz_max > 8 * nt ||
        a.rows_max > nt || (a.rows_max + nt / a.lanes - 1) / (nt / a.lanes) > (nt == 512 ? 2 : 4))
        return false;
    return a.slots < 65535 && ring_lds_bytes(a) <= 160 * 1024;
}

dspmm_status_t launch_ring(const RingLaunch &a, hipStream_t st) {
    if (a.nstreams == 0) return DSPMM_STATUS_SUCCESS;
    if (!ring_supported(a)) return DSPMM_STATUS_NOT_SUPPORTED;
    if (a.bf16) {
        switch (a.lanes) {
        case 1: return launch_ring_vl<uint16_t, 8, 1>(a, st);
        case 2: return launch_ring_vl<uint16_t, 8, 2>(a, st);
        case 4: return launch_ring_vl<uint16_t, 8, 4>(a, st);
        case 8: return launch_ring_vl<uint16_t, 8, 8>(a, st);
        case 16: return launch_ring_vl<uint16_t, 8, 16>(a, st);
        case 32: return launch_ring_vl<uint16_t, 8, 32>(a, st);
        default: return DSPMM_STATUS_NOT_SUPPORTED;
        }
    }
    switch (a.lanes) {
    case 1: return launch_ring_vl<float, 4, 1>(a, st);
    case 2: return launch_ring_vl<float, 4, 2>(a, st);
    case 4: return launch_ring_vl<float, 4, 4>(a, st);
    case 8: return launch_ring_vl<float, 4, 8>(a, st);
    case 16: return launch_ring_vl<float, 4, 16>(a, st);
    case 32: return launch_ring_vl<float, 4, 32>(a, st);
    default: return DSPMM_STATUS_NOT_SUPPORTED;
    }
}

}  // namespace dspmm
