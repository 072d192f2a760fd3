// ring_kernels.hip -- k_ring: row-major SpMM through a ring of B rows held
// in LDS and filled by a host-computed schedule (ring_plan.cpp).
//
// One workgroup walks one stream of tiles.  The plan says, per tile, which
// B rows to load into which ring slots before the tile runs, and gives
// every nonzero the slot of its B row.  The walk is software-pipelined two
// tiles deep with one barrier per tile.  In the iteration of tile k:
//   (A) tile k+1's B-row pieces and (slot, value) metadata, loaded into
//       registers one iteration earlier, are written to LDS -- the plan
//       never puts a fill of tile k+1 into a slot tile k reads, so other
//       waves may still be computing tile k; the metadata alternates
//       between two buffers;
//   (B) the loads for tile k+2 go out: its B-row pieces (addressed by
//       fill-list entries loaded one iteration earlier still), its
//       metadata, and tile k+3's fill-list entries;
//   (C) tile k is computed from LDS, its C rows kept in registers and
//       stored in (D), in the next iteration between (A) and (B), and
//   (E) a barrier ends the iteration.
// A load thus has a whole tile of compute to land before anything waits for
// it.  A group of L lanes x VW columns sums one row at a time in its nonzero
// order, with a separately rounded multiply and add, so every row equals the
// reference host loop (spmvspmm_benchmark.cu:11-39) bit for bit, with FP32
// or BF16-stored B.
//
// What this buys on a grid stencil: the row-gathering kernels fetch a B row
// through L1 for every nonzero (27 per output row on a 27-point stencil);
// here a workgroup's B traffic is its tiles' halos, about two B rows per
// output row, and the 27 reads per output row are LDS reads.
#include <algorithm>

#include "common.hpp"
#include "ring.hpp"

namespace dspmm {

__host__ __device__ inline size_t ring_bytes16(size_t b) { return (b + 15) / 16 * 16; }

// LDS layout: ring (slots + the zero slot) | two metadata buffers, each:
// ring byte offsets (int) and values (float) of a tile's nonzeros, 8 spare
// entries each, and the tile's row ids (rows_max int) | the stream's tile
// descriptors (int4) and fill offsets (int) | one tile's fill-list entries
// (fill_max + 1 int2, the last a harmless {column, slot} for lanes past
// the end).
struct RingLayout {
    size_t offs, vals, rows;      // within a metadata buffer
    size_t meta, mstride;         // first metadata buffer, distance to the second
    size_t tiles, foff, ents, total;
};
__host__ __device__ inline RingLayout ring_layout(int slots, int nv, bool bf16, int rows_max,
                                                  int nnz_max, int stiles, int fill_max) {
    RingLayout l;
    const size_t ring = ring_bytes16((size_t)(slots + 1) * ring_row_elems(nv, bf16) * (bf16 ? 2 : 4));
    l.offs = 0;
    l.vals = 4 * ((size_t)nnz_max + 8);
    l.rows = l.vals + 4 * ((size_t)nnz_max + 8);
    const size_t mb = ring_bytes16(l.rows + 4 * (size_t)rows_max);
    l.meta = ring;
    l.mstride = mb;
    l.tiles = ring + 2 * mb;
    l.foff = l.tiles + 16 * (size_t)stiles;
    l.ents = ring_bytes16(l.foff + 4 * ((size_t)stiles + 1));
    l.total = l.ents + 8 * ((size_t)fill_max + 1);
    return l;
}

size_t ring_lds_bytes(const RingLaunch &a) {
    return ring_layout(a.slots, a.nv, a.bf16, a.rows_max, a.nnz_max, a.stream_tiles_max,
                       a.fill_max).total;
}

namespace {

// Loads and stores at a uniform base plus a 32-bit byte offset (the base in
// scalar registers, one offset register per lane): the kernel is only taken
// when B, C and the plan arrays each span under 4 GiB and row indices stay
// under 2^24, so the offsets are exact 24 x 24-bit products.
template <typename T>
__device__ __forceinline__ const T &at32(const void *base, uint32_t off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const unsigned char *>(base) + off);
}

struct MetaBuf {
    int *offs;               // per nonzero of the tile: ring byte offset of its B row
    float *vals;             // per nonzero: its value
    int *rows;               // row id per row of the tile
};
__device__ __forceinline__ MetaBuf meta_buf(unsigned char *lds, const RingLayout &l, int b) {
    unsigned char *base = lds + l.meta + (size_t)b * l.mstride;
    return MetaBuf{reinterpret_cast<int *>(base + l.offs), reinterpret_cast<float *>(base + l.vals),
                   reinterpret_cast<int *>(base + l.rows)};
}

// A B row is `parts` 16-byte pieces (a power of two dividing the
// workgroup): thread t always moves piece q = t % parts, of fills
// t / parts, t / parts + NT / parts, ...  FR pieces per thread, fills
// [base, base + FR passes) of a tile whose fills are entries [g0, g0 + nf)
// of the plan's list; FR is picked at launch so that FR passes cover every
// tile after a stream's first (plans that need more are not taken), so the
// walk itself carries a tile's fills in registers in one go.  No branch
// surrounds a load:
// a thread past the end reloads entry g0, which exists (the plan keeps one
// entry past its last), and is marked not ok / dst = -1.
typedef unsigned int piece_t __attribute__((ext_vector_type(4)));
struct PieceMap {
    int q;                   // this thread's piece of every row it moves
    int f;                   // its first fill
    int step;                // fills per pass of the workgroup
};
template <int FR>
struct FillEnt {
    int2 e[FR];              // {column, slot}
    bool ok[FR];
};
template <int FR>
struct FillRegs {
    piece_t v[FR];
    int dst[FR];             // byte offset in the ring, -1: no piece
};

template <int FR>
__device__ __forceinline__ void fill_entries(const RingLaunch &a, const PieceMap &pm, int g0,
                                             int nf, int base, FillEnt<FR> &e) {
#pragma unroll
    for (int i = 0; i < FR; i++) {
        const int k = base + pm.f + i * pm.step;
        e.ok[i] = k < nf;
        e.e[i] = at32<int2>(a.fills, (uint32_t)(g0 + (k < nf ? k : 0)) * 8u);
    }
}

// The entries of a tile whose fills are issued one iteration later pass
// through LDS (each thread reads back what a thread of its own wave wrote:
// the threads of one fill are parts <= 64 consecutive lanes), so that the
// registers they arrive in are free again at once.
template <int FR>
__device__ __forceinline__ void ents_stash(int2 *ents, const PieceMap &pm, int fill_max,
                                           const FillEnt<FR> &e) {
#pragma unroll
    for (int i = 0; i < FR; i++) ents[min(pm.f + i * pm.step, fill_max)] = e.e[i];
}
template <int FR>
__device__ __forceinline__ void ents_fetch(const int2 *ents, const PieceMap &pm, int fill_max,
                                           int nf, FillEnt<FR> &e) {
#pragma unroll
    for (int i = 0; i < FR; i++) {
        const int k = pm.f + i * pm.step;
        e.ok[i] = k < nf;
        e.e[i] = ents[min(k, fill_max)];
    }
}

template <typename TB, int FR>
__device__ __forceinline__ void fill_issue(const RingLaunch &a, const PieceMap &pm,
                                           const FillEnt<FR> &e, int row_bytes, FillRegs<FR> &f) {
    const unsigned ldb_bytes = (unsigned)a.ldb * (unsigned)sizeof(TB);
#pragma unroll
    for (int i = 0; i < FR; i++) f.dst[i] = e.ok[i] ? e.e[i].y * row_bytes + pm.q * 16 : -1;
#pragma unroll
    for (int i = 0; i < FR; i++)
        f.v[i] = at32<piece_t>(a.B, __umul24((unsigned)e.e[i].x, ldb_bytes) + pm.q * 16);
}

template <int FR>
__device__ __forceinline__ void fill_store(unsigned char *ring, const FillRegs<FR> &f) {
#pragma unroll
    for (int i = 0; i < FR; i++)
        if (f.dst[i] >= 0) *reinterpret_cast<piece_t *>(ring + f.dst[i]) = f.v[i];
}

// A tile's fills from `base` on, synchronously (a stream's first tile).
template <typename TB, int FR>
__device__ __forceinline__ void fill_sync(const RingLaunch &a, unsigned char *ring,
                                          const PieceMap &pm, int g0, int nf, int base,
                                          int row_bytes) {
    for (int b = base; b < nf; b += pm.step * FR) {
        FillEnt<FR> e;
        FillRegs<FR> f;
        fill_entries<FR>(a, pm, g0, nf, b, e);
        fill_issue<TB, FR>(a, pm, e, row_bytes, f);
        fill_store<FR>(ring, f);
    }
}

// Metadata of one tile: thread t holds nonzeros [8t, 8t + 8) of the tile
// (values and slots -- one byte each where the plan's slots fit a byte,
// else two -- aligned in the plan) and row t.  Waves wholly
// past the tile's nonzeros (rows) load nothing; within a wave a thread past
// the end reloads the last chunk (row).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
struct MetaRegs {
    f32x4_t v0, v1;
    piece_t sl;              // 8 slots: bytes of sl.x, sl.y, or 16-bit halves of all four
    int row;
    int nch, nrows, lmax;
};

__device__ __forceinline__ void meta_load(const RingLaunch &a, int4 td, MetaRegs &m) {
    const int t = threadIdx.x;
    // the descriptor is the same in every lane (read from LDS): scalar
    td.x = __builtin_amdgcn_readfirstlane(td.x);
    td.y = __builtin_amdgcn_readfirstlane(td.y);
    td.z = __builtin_amdgcn_readfirstlane(td.z);
    td.w = __builtin_amdgcn_readfirstlane(td.w);
    // td.y: the tile's longest row; its padded entries are rows x that
    // length rounded up to 4, then to 8
    m.nrows = td.w - td.z;
    m.lmax = td.y;
    m.nch = (m.nrows * ((td.y + 3) & ~3) + 7) >> 3;
    const int w0 = __builtin_amdgcn_readfirstlane(t) & ~63;   // the wave's first thread
    if (w0 < m.nch) {
        const int c = min(t, m.nch - 1);
        const uint32_t v = (uint32_t)(td.x + 8 * c) * 4u, sl = (uint32_t)(td.x + 8 * c) * 2u;
        m.v0 = at32<f32x4_t>(a.pval, v);
        m.v1 = at32<f32x4_t>(a.pval, v + 16);
        // two 8-byte loads either way (one-byte slots: the same eight bytes
        // twice), so that no branch surrounds a load
        const uint32_t s0 = a.slot_bytes == 1 ? sl >> 1 : sl, s1 = a.slot_bytes == 1 ? s0 : s0 + 8;
        const uint2 lo = at32<uint2>(a.pslot, s0), hi = at32<uint2>(a.pslot, s1);
        m.sl = piece_t{lo.x, lo.y, hi.x, hi.y};
    }
    if (w0 < m.nrows) m.row = at32<int>(a.row_info, (uint32_t)(td.z + min(t, m.nrows - 1)) * 4u);
}

__device__ __forceinline__ void meta_store(const MetaBuf &s, const MetaRegs &m, int row_bytes,
                                           int slot_bytes) {
    const int t = threadIdx.x;
    if (t < m.nch) {
        // slot (< 2^16) x row bytes (< 2^16): a 24-bit multiply is exact
        const unsigned rb = (unsigned)row_bytes;
        int4 *o = reinterpret_cast<int4 *>(s.offs + 8 * t);
        if (slot_bytes == 1) {
            o[0] = make_int4(__umul24(m.sl[0] & 0xffu, rb), __umul24((m.sl[0] >> 8) & 0xffu, rb),
                             __umul24((m.sl[0] >> 16) & 0xffu, rb), __umul24(m.sl[0] >> 24, rb));
            o[1] = make_int4(__umul24(m.sl[1] & 0xffu, rb), __umul24((m.sl[1] >> 8) & 0xffu, rb),
                             __umul24((m.sl[1] >> 16) & 0xffu, rb), __umul24(m.sl[1] >> 24, rb));
        } else {
            o[0] = make_int4(__umul24(m.sl[0] & 0xffffu, rb), __umul24(m.sl[0] >> 16, rb),
                             __umul24(m.sl[1] & 0xffffu, rb), __umul24(m.sl[1] >> 16, rb));
            o[1] = make_int4(__umul24(m.sl[2] & 0xffffu, rb), __umul24(m.sl[2] >> 16, rb),
                             __umul24(m.sl[3] & 0xffffu, rb), __umul24(m.sl[3] >> 16, rb));
        }
        f32x4_t *v = reinterpret_cast<f32x4_t *>(s.vals + 8 * t);
        v[0] = m.v0;
        v[1] = m.v1;
    }
    if (t < m.nrows) s.rows[t] = m.row;
}

// Group g sums rows g, g + G, ..., g + (R - 1) G of the tile side by side
// into registers.  A row past the tile's end recomputes the tile's last
// row and lanes past nv the row's last VW columns, and store_rows writes
// those again, the same bits to the same place: every lane issues the same
// stores, with no branch, so the count of memory operations in flight
// after them is known and a later wait for loads does not wait for them.
// All rows of a tile are padded to one length, the longest row rounded up
// to 4 (zero-slot entries), so the loop over a row's entries, four at a
// time (offsets and values in 16-byte reads), has no bounds inside; it ends
// at the longest row, so at most the shorter rows' padding is summed.
template <int VW, int R>
struct TileRows {
    float acc[R][VW];
    uint32_t dst[R];         // byte offset in C of the row's VW columns
};

// One step of four entries of every row the group holds (n of them summed:
// the rest of a row's last step past the tile's longest row are padding,
// whose +0 products are skipped -- the same sums).
template <typename TB, int VW, int R, int N>
__device__ __forceinline__ void ring_step(const unsigned char *rbase, const MetaBuf &s,
                                          const int (&base)[R], int p, TileRows<VW, R> &out) {
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
        if (N > 1) load_b<TB, VW>(b[j][1], reinterpret_cast<const TB *>(rbase + o[j].y));
        if (N > 2) load_b<TB, VW>(b[j][2], reinterpret_cast<const TB *>(rbase + o[j].z));
        if (N > 3) load_b<TB, VW>(b[j][3], reinterpret_cast<const TB *>(rbase + o[j].w));
    }
#pragma unroll
    for (int j = 0; j < R; j++)
#pragma unroll
        for (int k = 0; k < VW; k++) {
            float c = out.acc[j][k];
            c = c + v[j].x * b[j][0][k];
            if (N > 1) c = c + v[j].y * b[j][1][k];
            if (N > 2) c = c + v[j].z * b[j][2][k];
            if (N > 3) c = c + v[j].w * b[j][3][k];
            out.acc[j][k] = c;
        }
}

template <typename TB, int VW, int L, int R, int NT>
__device__ __forceinline__ void compute_tile(const RingLaunch &a, const unsigned char *ring,
                                             const MetaBuf &s, int nrows, int lmax,
                                             TileRows<VW, R> &out) {
    constexpr int G = NT / L;
    const int t = threadIdx.x, g = t / L;
    const int col0 = (t % L) * VW;
    // lanes past nv read (and write) a valid column of the row (nv is a
    // multiple of VW)
    const int colc = min(col0, a.nv - VW);
    const unsigned char *rbase = ring + colc * (int)sizeof(TB);
    const int len = (lmax + 3) & ~3;     // every row's padded length
    int base[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int ie = min(g + j * G, nrows - 1);
        base[j] = (int)__umul24((unsigned)ie, (unsigned)len);   // both under 2^24
        out.dst[j] = __umul24((unsigned)s.rows[ie], (unsigned)a.ldc * 4u) + colc * 4;
#pragma unroll
        for (int k = 0; k < VW; k++) out.acc[j][k] = 0.f;
    }
    // whole steps of four, then the entries of the longest row past them
    // (lmax is wave-uniform: the branches below are scalar)
    const int full = lmax & ~3;
    for (int p = 0; p < full; p += 4) ring_step<TB, VW, R, 4>(rbase, s, base, p, out);
    switch (lmax - full) {
    case 1: ring_step<TB, VW, R, 1>(rbase, s, base, full, out); break;
    case 2: ring_step<TB, VW, R, 2>(rbase, s, base, full, out); break;
    case 3: ring_step<TB, VW, R, 3>(rbase, s, base, full, out); break;
    default: break;
    }
}

template <int VW, int R>
__device__ __forceinline__ void store_rows(const RingLaunch &a, const TileRows<VW, R> &r) {
#pragma unroll
    for (int j = 0; j < R; j++)
        store_vec_nt<VW>(reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(a.C) + r.dst[j]),
                         r.acc[j]);
}

}  // namespace

// Workgroups of NT = 256 or 512 threads.  The ring takes a large part of a
// CU's LDS, so few workgroups share a CU: registers are budgeted for two
// (NT / 128 waves per SIMD), keeping the loads in flight in registers
// across the compute loop.
template <typename TB, int VW, int L, int R, int NT, int FR>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, NT / 128)))
void k_ring(RingLaunch a) {
    extern __shared__ __align__(16) unsigned char rsm[];
    const int row_bytes = ring_row_elems(a.nv, sizeof(TB) == 2) * (int)sizeof(TB);
    const RingLayout lay = ring_layout(a.slots, a.nv, sizeof(TB) == 2, a.rows_max, a.nnz_max,
                                       a.stream_tiles_max, a.fill_max);
    int4 *tiles = reinterpret_cast<int4 *>(rsm + lay.tiles);
    int *foff = reinterpret_cast<int *>(rsm + lay.foff);
    int2 *ents = reinterpret_cast<int2 *>(rsm + lay.ents);
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
    const int f0 = a.fill_off[t0];
    {   // the stream's tile descriptors and fill offsets, and the zero slot
        // (which padding entries read), into LDS
        for (int i = t; i < ns; i += NT) {
            tiles[i] = a.tiles[t0 + i];
            foff[i] = a.fill_off[t0 + i] - f0;
        }
        if (t == 0) {
            foff[ns] = a.fill_off[t1] - f0;
            ents[a.fill_max] = make_int2(0, 0);
        }
        for (int i = t; i < row_bytes / 4; i += NT)
            reinterpret_cast<int *>(rsm + (size_t)a.slots * row_bytes)[i] = 0;
        __syncthreads();
    }
    // fills of tile i: entries [f0 + fo(i), f0 + fo(i + 1)) of the plan's list
    auto fo = [&](int i) { return __builtin_amdgcn_readfirstlane(foff[i]); };
    MetaRegs m;
    FillEnt<FR> ent, cur;
    FillRegs<FR> f;
    int nrows, lmax;
    {   // tile 0, synchronously; tile 1's loads in flight; tile 2's entries
        meta_load(a, tiles[0], m);
        fill_sync<TB, FR>(a, rsm, pm, f0, fo(1), 0, row_bytes);
        meta_store(meta_buf(rsm, lay, 0), m, row_bytes, a.slot_bytes);
        nrows = m.nrows;
        lmax = m.lmax;
        // (for a one- or two-tile stream: its last tile's, never stored)
        const int k1 = min(1, ns - 1), k2 = min(2, ns - 1);
        fill_entries<FR>(a, pm, f0 + fo(k1), fo(k1 + 1) - fo(k1), 0, ent);
        fill_issue<TB, FR>(a, pm, ent, row_bytes, f);
        meta_load(a, tiles[k1], m);
        fill_entries<FR>(a, pm, f0 + fo(k2), fo(k2 + 1) - fo(k2), 0, ent);
        __syncthreads();
    }
    TileRows<VW, R> out;                  // the previous tile's rows
    for (int k = 0; k < ns; k++) {
        int nrows_next = 0, lmax_next = 0;
        // (A) and (B) at a raised wave priority, (C) at the default: the
        // SIMD's arbiter then issues a wave's stores and loads ahead of other
        // waves' multiplies, so the loads for later tiles go out sooner
        // (BF16 B, VALU-bound: 176 -> 169 us on the 100^3 stencil; FP32 B
        // within 1%)
        __builtin_amdgcn_s_setprio(3);
        if (k + 1 < ns) {
            // (A) tile k + 1 into LDS: its slots are none that tile k reads,
            // and tile k - 1, whose metadata buffer it takes, is done
            fill_store<FR>(rsm, f);
            meta_store(meta_buf(rsm, lay, (k + 1) & 1), m, row_bytes, a.slot_bytes);
            nrows_next = m.nrows;
            lmax_next = m.lmax;
        }
        // (B) tile k + 2's loads, tile k + 3's fill-list entries.  Issued on
        // every iteration, for the stream's last tile past its end (loaded,
        // never stored), so that no branch surrounds them.
        const int k2 = min(k + 2, ns - 1), k3 = min(k + 3, ns - 1);
        ents_stash<FR>(ents, pm, a.fill_max, ent);
        // (D) tile k - 1's C rows go out here, ahead of this iteration's
        // loads: a later wait for those loads then finds the stores long done
        if (k > 0) store_rows<VW, R>(a, out);
        // (the compiler must read the entries back, not reuse the registers)
        asm volatile("" ::: "memory");
        ents_fetch<FR>(ents, pm, a.fill_max, fo(k2 + 1) - fo(k2), cur);
        fill_issue<TB, FR>(a, pm, cur, row_bytes, f);
        meta_load(a, tiles[k2], m);
        fill_entries<FR>(a, pm, f0 + fo(k3), fo(k3 + 1) - fo(k3), 0, ent);
        // (C) tile k
        __builtin_amdgcn_s_setprio(0);
        compute_tile<TB, VW, L, R, NT>(a, rsm, meta_buf(rsm, lay, k & 1), nrows,
                                       __builtin_amdgcn_readfirstlane(lmax), out);
        nrows = nrows_next;
        lmax = lmax_next;
        __syncthreads();                  // (E)
    }
    store_rows<VW, R>(a, out);            // the last tile's rows
}

// Workgroup size: 512 threads where the LDS a workgroup needs leaves room
// for at most two per CU (more waves each), else 256 (more workgroups).
static int ring_threads(const RingLaunch &a) {
    if (ring_lds_bytes(a) * 3 > 160 * 1024 && (a.lanes == 8 || a.lanes == 16 || a.lanes == 32) &&
        (a.rows_max + 512 / a.lanes - 1) / (512 / a.lanes) <= 2 && a.nnz_max <= 8 * 512)
        return 512;
    return 256;
}

// Passes of the workgroup that carry a tile's fills (a stream's first tile
// aside): 2, 4 or 8 pieces per thread; 0 when a tile needs more.
static int ring_fill_passes(const RingLaunch &a, int nt) {
    const int per_pass = nt / (a.nv * (a.bf16 ? 2 : 4) / 16);
    for (int fr : {2, 4, 8})
        if (a.fill_max <= fr * per_pass) return fr;
    return 0;
}

namespace {
template <typename TB, int VW, int L, int R, int NT>
void launch_ring_fr(const RingLaunch &a, size_t lds, hipStream_t st) {
    switch (ring_fill_passes(a, NT)) {
    case 2: hipLaunchKernelGGL((k_ring<TB, VW, L, R, NT, 2>), dim3(a.nstreams), dim3(NT), lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_ring<TB, VW, L, R, NT, 4>), dim3(a.nstreams), dim3(NT), lds, st, a); break;
    default: hipLaunchKernelGGL((k_ring<TB, VW, L, R, NT, 8>), dim3(a.nstreams), dim3(NT), lds, st, a); break;
    }
}
template <typename TB, int VW, int L, int NT>
dspmm_status_t launch_ring_n(const RingLaunch &a, hipStream_t st) {
    constexpr int G = NT / L;
    const size_t lds = ring_lds_bytes(a);
    const int rows_per_group = (a.rows_max + G - 1) / G;
    if (rows_per_group <= 1)
        launch_ring_fr<TB, VW, L, 1, NT>(a, lds, st);
    else if (rows_per_group == 2 || NT == 512)
        launch_ring_fr<TB, VW, L, 2, NT>(a, lds, st);
    else
        launch_ring_fr<TB, VW, L, (NT == 512 ? 2 : 4), NT>(a, lds, st);
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
        parts < 1 || parts > 64 || (parts & (parts - 1)) || (a.ldb * eb) % 16 ||
        !aligned(a.B, 16) || !aligned(a.C, 16) || (a.ldc % a.vw) || a.nnz_max > 8 * nt ||
        a.rows_max > nt || (a.rows_max + nt / a.lanes - 1) / (nt / a.lanes) > (nt == 512 ? 2 : 4) ||
        ring_fill_passes(a, nt) == 0)
        return false;
    // 32-bit offsets: B and C under 4 GiB, row indices under 2^24 (exact
    // 24-bit products), the plan's arrays under 4 GiB
    constexpr int64_t k4G = (int64_t)1 << 32, k16M = (int64_t)1 << 24;
    if (a.nc >= k16M || a.nr >= k16M || (int64_t)a.nc * a.ldb * eb >= k4G ||
        (int64_t)a.nr * a.ldc * 4 >= k4G || a.ldb * eb >= k16M || a.ldc * 4 >= k16M ||
        a.plan_entries * 4 >= k4G || a.plan_fills * 8 >= k4G)
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
