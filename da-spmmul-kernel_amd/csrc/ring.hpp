// ring.hpp -- the B-row ring kernel (k_ring): a software-managed cache of B
// rows in LDS, filled by a host-computed schedule.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "internal.hpp"

namespace dspmm {

// Tile shape of a ring plan.  With 1 <= s_mid < s_big (s_big a multiple of
// s_mid, s_mid >= w) a tile is w consecutive rows in each of h runs s_mid
// apart, and a stream visits the same block in `zchunk` successive
// s_big-planes; otherwise a tile is w * h consecutive rows and a stream
// `zchunk` consecutive tiles.  `slots` is the ring's capacity in B rows
// (0: the largest tile's distinct columns, rounded up to 8).
struct RingShape {
    int w = 8, h = 4;
    int s_mid = 1, s_big = 1;
    int zchunk = 25;
    int slots = 0;
};

// Host plan (see ring_plan.cpp).
struct RingPlanHost {
    // {p_begin, longest row, q_begin, q_end} (q: index into row_info).  The tile's
    // nonzeros are p_begin + [0, rows x padded length), the padded length being
    // the longest row rounded up to 4, then rounded up to 8 entries.
    std::vector<int4> tiles;
    std::vector<int> stream_off;      // nstreams + 1 offsets into tiles
    std::vector<int2> row_info;       // {row, end of its padded nonzeros relative to p_begin}
    std::vector<int> fill_off;        // ntiles + 1 offsets into fills
    std::vector<int2> fills;          // {column, slot} loaded before a tile runs
    std::vector<uint16_t> pslot;      // per nonzero (tile-row order): the slot of its column
    std::vector<float> pval;          // per nonzero: its value (0 for padding)
    int slots = 0;                    // ring capacity the plan was made for
    int zero_slot = 0;                // = slots: the all-zero slot padding entries point at
    int rows_max = 0;                 // rows of the largest tile
    int nnz_max = 0;                  // padded nonzeros of the largest tile
    int fill_max = 0;                 // fills of the largest tile that is not a stream's first
    int stream_tiles_max = 0;         // tiles of the longest stream
    int stream_fills_max = 0;         // fills of the stream with the most
    int64_t total_fills = 0;
    int64_t pad_entries = 0;          // entries of all tiles' rows, padding included
};

// Tiles per stream (zchunk) for a GPU of `cus` CUs.  The walks of a plan
// spread over the CUs, and the walks a CU holds share its throughput, so the
// kernel takes about ceil(walks / cus) x (kRingFirstTileCost + tiles per
// walk) tile times: a walk's first tile costs about kRingFirstTileCost
// tiles more, as its B rows (a stencil's three halo planes) all load before
// it runs, with nothing to overlap them.  Returns the length, at most
// kRingMaxChunk, with the least such cost, the longest on ties.  E.g. 325
// walks of 100 planes on 256 CUs: 34 (975 walks, 4 per CU, 4 x 36 = 144
// tile times) -- 10 takes 13 x 12 = 156, 100 takes 2 x 102.
constexpr int kRingFirstTileCost = 2;
constexpr int kRingMaxChunk = 256;
int ring_auto_chunk(int nr, const RingShape &sh, int cus);

// False when a tile needs more distinct columns than the ring holds (or the
// shape is invalid).
bool build_ring_plan(const std::vector<int> &rp, const std::vector<int> &ci,
                     const std::vector<float> &va, int nr, int nc, const RingShape &sh,
                     RingPlanHost &p);

// One launch of k_ring: one workgroup per stream.
struct RingLaunch {
    const int4 *tiles = nullptr;
    const int *stream_off = nullptr;
    const int *row_info = nullptr;    // row ids of the plan's tiles, tile after tile
    const int *fill_off = nullptr;
    const int2 *fills = nullptr;
    const void *pslot = nullptr;      // slot per entry, slot_bytes wide
    int slot_bytes = 2;               // 1 when every slot (the zero slot too) is under 256
    const float *pval = nullptr;
    const void *B = nullptr;
    float *C = nullptr;
    int64_t ldb = 0, ldc = 0;
    int nv = 0;
    int nstreams = 0;
    int slots = 0;
    int rows_max = 0, nnz_max = 0;
    int stream_tiles_max = 0, stream_fills_max = 0;
    int fill_max = 0;                 // fills of the largest tile after a stream's first
    int nr = 0, nc = 0;               // rows of C and of B
    int64_t plan_entries = 0;         // entries of pslot / pval
    int64_t plan_fills = 0;           // entries of fills
    int vw = 4, lanes = 16;
    bool bf16 = false;
};
// LDS bytes of one k_ring workgroup: the ring, two tiles' metadata and
// rows, and its stream's tile descriptors and fill offsets.
size_t ring_lds_bytes(const RingLaunch &a);
// Elements per ring row: nv, plus a 16-byte pad for rows under 128 bytes
// (there lane groups of under eight lanes share a 16-byte-per-lane LDS
// read, and rows a power of two apart would meet in the same banks).
__host__ __device__ inline int ring_row_elems(int nv, bool bf16) {
    const int eb = bf16 ? 2 : 4;
    return nv + (nv * eb >= 128 ? 0 : 16 / eb);
}
// Whether these operands and this plan fit the kernel (one slab, 16-byte
// pieces, a tile's metadata in one register load, the LDS budget).
bool ring_supported(const RingLaunch &a);
dspmm_status_t launch_ring(const RingLaunch &a, hipStream_t st);

}  // namespace dspmm
