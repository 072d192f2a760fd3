// matrix.hpp -- the device-resident CSR handle and its nnz-tile plans
// (shared by the single-matrix and batched native paths).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.hpp"
#include "ring.hpp"

namespace dspmm {

struct NbPlan {
    int S = 0;
    int ntiles = 0, nsplits = 0, nslots = 0;
    int4 *d_tiles = nullptr;
    int4 *d_splits = nullptr;         // {0, row, first_slot, nchunks}
    std::vector<int4> h_tiles, h_splits;
    size_t bytes = 0;
    // column-cached plans only: per-tile distinct columns, per-nonzero position
    int *d_ucols = nullptr;
    uint16_t *d_lidx = nullptr;
    double reuse = 0;                 // nonzeros per B row read from L2
    // row-blocked plans only: per-block list offsets, {column, row mask}, values
    int *d_boff = nullptr;
    int2 *d_list = nullptr;
    float *d_vals = nullptr;
    bool usable = true;               // false: rows unsorted or with repeated columns
};

// A B-row ring plan (k_ring) uploaded to the device.
struct RingPlan {
    int4 *d_tiles = nullptr;
    int *d_streams = nullptr;
    int *d_rows = nullptr;           // row ids, tile after tile
    int *d_fill_off = nullptr;
    int2 *d_fills = nullptr;
    void *d_pslot = nullptr;         // slot per entry: 1 byte when slots < 256, else 2
    int slot_bytes = 2;
    float *d_pval = nullptr;
    int nstreams = 0, ntiles = 0, slots = 0, rows_max = 0, nnz_max = 0;
    int stream_tiles_max = 0, stream_fills_max = 0, fill_max = 0;
    int64_t fills = 0;
    int64_t entries = 0, fill_entries = 0;   // lengths of d_pslot / d_pval and of d_fills
    size_t bytes = 0;
    bool usable = false;     // false: a tile needs more B rows than the ring holds
};

// LDS budget of one nnz-tiled workgroup (keeps two workgroups per CU).
constexpr size_t kMaxTileLds = 64 * 1024;

void build_tiles(const std::vector<int> &rp, int nr, int S, std::vector<int4> &tiles,
                 std::vector<int4> &splits, int &nslots);
// SpMV plan: tiles of <= T nonzeros and <= 256 rows, each row <= row_max
// long; a longer row gets its own tile(s) of <= T nonzeros.
void build_spmv_tiles(const std::vector<int> &rp, int nr, int T, int row_max,
                      std::vector<int4> &tiles, std::vector<int4> &splits, int &nslots);

// Column-cached plan: tiles of whole rows with at most umax distinct columns,
// at most smax nonzeros and kCachedRowsMax rows; a row that does not fit alone
// becomes a direct tile {row, -1, p_begin, u}.  tiles has a sentinel entry
// whose .w is the total number of listed columns.
void build_cached_tiles(const std::vector<int> &rp, const std::vector<int> &ci, int nr, int nc,
                        int umax, int smax, std::vector<int4> &tiles, std::vector<int> &ucols,
                        std::vector<uint16_t> &lidx, double &reuse);

// Row-blocked plan: blocks of R rows `stride` apart (block b holds rows
// (b / stride) * R * stride + b % stride + i * stride; blocks past the last
// row are empty; false if the block count overflows an int); per block the ascending
// distinct columns of its rows, each with the mask of rows that have it, and
// per listed column R values (the rows' values, 0 where the mask bit is
// clear).  False if a row is unsorted or repeats a column (the plan then
// cannot keep every row's nonzero order).
bool build_row_blocks(const std::vector<int> &rp, const std::vector<int> &ci,
                      const std::vector<float> &va, int nr, int R, int stride,
                      std::vector<int> &boff, std::vector<int2> &list, std::vector<float> &vals);

enum PlanKind { kPlanNnzTiled = 0, kPlanSpmv = 1, kPlanCached = 2, kPlanRowBlock = 3 };
constexpr size_t kMaxCachedLds = 80 * 1024;   // two column-cached workgroups per CU

}  // namespace dspmm

struct dspmm_matrix_s {
    int nr = 0, nc = 0, nnz = 0;
    int *rowPtr = nullptr, *rowIdx = nullptr, *colIdx = nullptr;
    float *values = nullptr;
    // Host copies kept from creation: plans and sampled features are built
    // from them, so no launch ever copies the matrix back from the device.
    std::vector<int> h_rowPtr, h_colIdx;
    std::vector<float> h_values;
    int max_len = 0;
    double mean = 0, cv = 0;
    size_t base_bytes = 0;
    std::mutex mu;
    std::map<int64_t, dspmm::NbPlan> plans;                        // by 4 * (tile size + 16 * (stride - 1)) + kind
    double reuse16 = -1;     // sampled nonzeros per distinct column over 16-row windows (-1: not yet)
    int rb_stride = -1;      // row distance of the most shared columns, sampled (-1: not yet)
    int rb_stride2 = -1;     // the best distance below it that divides it (1: none)
    // ring plans by (w, h, tiles per stream)
    std::map<std::tuple<int, int, int>, dspmm::RingPlan> ring_plans;
    // tiles per stream chosen for tile shape (w, h) when the config leaves it to the library
    std::map<std::pair<int, int>, int> ring_auto;
    std::map<std::tuple<int, int, int>, dspmm_config> tuned;       // (layout, nv, btype)
    std::map<std::tuple<int, int, int>, dspmm_config> prepared;    // AUTO's pick, by dspmm_matrix_prepare
    float *carry = nullptr;
    size_t carry_bytes = 0;
    void *tbuf = nullptr;       // column-major operands transposed for the row-major kernels
    size_t tbuf_bytes = 0;

    size_t hbm_bytes() {
        size_t b = base_bytes + carry_bytes + tbuf_bytes;
        if (rowIdx) b += sizeof(int) * (size_t)nnz;
        for (auto &kv : plans) b += kv.second.bytes;
        for (auto &kv : ring_plans) b += kv.second.bytes;
        return b;
    }
};

namespace dspmm {

// All of these expect the caller to hold A->mu.
// The plan of (S, kind, stride), built and uploaded on first use; with
// build == false a plan not built yet is DSPMM_STATUS_NOT_PREPARED.
dspmm_status_t get_plan(dspmm_matrix_s *A, int S, NbPlan **out, int kind = kPlanNnzTiled,
                        int stride = 1, bool build = true);
// Column reuse of neighbouring rows, sampled (computed once per matrix).
double column_reuse(dspmm_matrix_s *A);
// The row distance >= 2 at which up to 256 sampled rows share the most
// columns with the row that far below them (candidates: the 16 most
// frequent column-minus-row offsets; ties go to the larger distance), or 1
// when they share fewer than max(1, mean row length / 4) on average.
// Computed once per matrix.
int shared_row_stride(dspmm_matrix_s *A);
// The distance below shared_row_stride(A) that divides it and at which the
// sampled rows share the most columns (a stencil's line length below its
// plane length), or 1 when none qualifies.  Computed with it.
int shared_row_stride2(dspmm_matrix_s *A);
// The ring plan of tile shape (w, h) and `zchunk` tiles per stream, built
// on first use (build == false: DSPMM_STATUS_NOT_PREPARED if missing).
dspmm_status_t get_ring_plan(dspmm_matrix_s *A, int w, int h, int zchunk, RingPlan **out,
                             bool build = true);
// The plan's shape for tile w x h (one row distance: w * h consecutive rows
// per tile, walked along that distance).  Reads the sampled row distances.
RingShape ring_shape(dspmm_matrix_s *A, int w, int h, int zchunk);
// A launch's plan arrays and sizes from `plan`.
void ring_bind_plan(dspmm_matrix_s *A, const RingPlan &plan, RingLaunch &a);
// Tiles per stream of a ring launch on tile shape (w, h) when the config
// leaves it to the library: ring_auto_chunk for the current GPU's CU count,
// cached per shape.
int ring_auto_z(dspmm_matrix_s *A, int w, int h);
int cached_smax(int umax);
int64_t carry_ld(int nv);
void complete_nb(dspmm_config &c, int nv, int64_t ldb, int64_t ldc, const void *B, const void *C,
                 bool bf16);
// The configuration a call runs: cfg unless AUTO, else the tuned one cached
// for (layout, nv, bt), else (with `learned`, row-major only) the
// nearest-neighbour selector's pick -- with BF16 B only a B-row ring pick --
// else the static heuristic.
dspmm_config resolve(dspmm_matrix_s *A, DenseLayout layout, int nv, dspmm_dtype_t bt,
                     const dspmm_config *cfg, bool learned = true);
// The static default (no timing, no corpus).
dspmm_config heuristic(const dspmm_matrix_s *A, DenseLayout layout, int nv, dspmm_dtype_t bt);
// Nearest-neighbour selection over the measured corpus (selector.cpp).
bool learned_select(dspmm_matrix_s *A, int nv, dspmm_config *out);

}  // namespace dspmm
