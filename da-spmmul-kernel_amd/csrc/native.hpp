// native.hpp -- host/device interface of the native kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace dspmm {

// One SpMM of the nnz-tiled path: a matrix's tile plan and its operands.
// Passed by value to the single-matrix kernels and as a device table to
// the batched kernels.
struct ItemDesc {
    const int4 *tiles;     // {row_begin, row_end | -1-slot, p_begin, p_end}
    const int *rowPtr;
    const int *colIdx;
    const float *values;
    const void *B;         // float or BF16 (uint16)
    float *C;
    float *carry;          // split-row partial sums: slots x ldcarry floats
    int64_t ldb, ldc, ldcarry;
    int nv;
    int slab_w;            // columns per group pass = lanes * vw
    int S;                 // tile size of the plan (max nonzeros / rows per tile)
    int cls;               // 8 * log2(vw) + log2(lanes)
    int ntiles;
    int nslabs;
    const int4 *splits;    // the plan's split records {0, row, first_slot, nchunks}
    int nsplits;
};

// One launch of the column-cached kernel (k_nnz_cached).  Tiles are
// {row_begin, row_end, p_begin, u_begin}; tiles[ntiles].w is the total
// number of cached columns.  A tile with row_end < 0 holds the single row
// row_begin, whose columns exceed the cache: it gathers B directly.
struct CachedLaunch {
    const int4 *tiles = nullptr;
    int ntiles = 0;
    const int *ucols = nullptr;        // the distinct columns of each tile, tile after tile
    const uint16_t *lidx = nullptr;    // per nonzero: position of its column in its tile's list
    const int *rowPtr = nullptr;
    const int *colIdx = nullptr;
    const float *values = nullptr;
    const void *B = nullptr;
    float *C = nullptr;
    int64_t ldb = 0, ldc = 0;
    int nv = 0;
    int umax = 0;                      // cached B rows per tile (LDS rows)
    int smax = 0;                      // nonzeros per tile
    int vw = 0, lanes = 0;
    int threads = 256;                 // workgroup size: 64, 128 or 256
    bool bf16 = false;
};
dspmm_status_t launch_nnz_cached(const CachedLaunch &a, hipStream_t st);
// LDS bytes of one k_nnz_cached workgroup.
size_t nnz_cached_lds_bytes(int umax, int smax, int nv, int vw, bool bf16);
constexpr int kCachedRowsMax = 256;    // rows per cached tile

// One launch of the row-blocked kernel (k_row_blocked): blocks of `rows`
// rows `stride` apart; list[boff[b] .. boff[b+1]) = {column, row mask} of block
// b in ascending column order; vals[u * rows + i] = row i's value at listed
// column u (0 where the mask bit is clear).
struct RowBlockLaunch {
    const int *boff = nullptr;
    const int2 *list = nullptr;
    const float *vals = nullptr;
    const int *rowPtr = nullptr;
    const void *B = nullptr;
    float *C = nullptr;
    int64_t ldb = 0, ldc = 0;
    int nr = 0, nv = 0, nblocks = 0;
    int rows = 4;                      // R: 2, 4 or 8
    int stride = 1;                    // row distance within a block
    int vw = 0, lanes = 0, slab_w = 0, nslabs = 1, unroll = 8;
    bool bf16 = false;
};
dspmm_status_t launch_row_blocked(const RowBlockLaunch &a, hipStream_t st);

// Arguments of one row-sequential launch (a lane group per row, indices
// staged per wavefront in LDS, `unroll` gathers in flight).
struct RowSeqLaunch {
    const int *rowPtr = nullptr;
    const int *colIdx = nullptr;
    const float *values = nullptr;
    const void *B = nullptr;
    float *C = nullptr;
    int64_t ldb = 0, ldc = 0;
    int nr = 0, nv = 0;
    int vw = 1, lanes = 1, slab_w = 1, nslabs = 1, unroll = 4;
    bool sliced = false;               // per-row slices of the staging windows
    bool bf16 = false;
};
bool row_seq_unroll_ok(int u);
// Pairs a row-sequential wavefront stages per window.
constexpr int kRowSeqWindow = 512;
dspmm_status_t launch_row_seq(const RowSeqLaunch &a, hipStream_t st);

// Arguments of one single-matrix nnz-tiled launch (plus its split combine).
struct NbLaunch {
    ItemDesc item{};
    int ntiles = 0;
    const int4 *splits = nullptr;  // {0, row, first_slot, nchunks}
    int nsplits = 0;
    bool bf16 = false;
    int vw = 1;
    int lanes = 1;
    int nslabs = 1;
    int unroll = 8;
    hipEvent_t mid_event = nullptr;  // recorded between the main kernel and the combine
};

// Arguments of one batched launch.
struct BatchLaunch {
    const ItemDesc *items = nullptr;   // device table
    const int4 *segs = nullptr;        // per-XCD work segments
    const int *xoff = nullptr;         // 9 segment offsets + 8 unit counts
    int lds_ints_per_wave = 0;
    int nblocks = 0;
    const int *split_prefix = nullptr; // count+1 prefix sums of the items' nsplits
    int count = 0;
    int nsplits = 0;                   // total
    bool bf16 = false;
    hipEvent_t mid_event = nullptr;
};

dspmm_status_t launch_nnz_tiled(const NbLaunch &a, hipStream_t st);
dspmm_status_t launch_nnz_tiled_batched(const BatchLaunch &a, hipStream_t st);
// SpMV (nv == 1) tiles: item.S is the tile size T, item.tiles a plan made
// by build_spmv_tiles.
dspmm_status_t launch_spmv_tiled(const NbLaunch &a, hipStream_t st);
constexpr int kSpmvRowMax = 256;   // longest row of a multi-row SpMV tile
constexpr int kSpmvThreadRow = 16; // longest row summed by one thread
constexpr int kSpmvLongTiles = 16; // rows up to 16 T nonzeros get one workgroup
bool spmv_tile_ok(int T);          // SpMV tile sizes with a kernel: 512 .. 4096
size_t nnz_tiled_lds_bytes(int lanes, int S);
int item_class(int vw, int lanes);
dspmm_status_t launch_fill_zero(void *p, size_t bytes, hipStream_t st);
// dst[c * ldd + r] = src[r * lds + c] for r < rows, c < cols (2- or 4-byte elements).
dspmm_status_t launch_transpose(const void *src, int64_t lds, int rows, int cols, void *dst,
                                int64_t ldd, int elem_bytes, hipStream_t st);
// Seeded uniform [-1, 1) fill of n FP32 (or BF16) values.
dspmm_status_t launch_fill_uniform(void *out, int64_t n, uint64_t seed, bool bf16, hipStream_t st);
dspmm_status_t launch_f32_to_bf16(const float *src, uint16_t *dst, int64_t n, hipStream_t st);

}  // namespace dspmm
