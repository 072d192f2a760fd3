// batch_schedule.hpp -- host-side schedule of a batched nnz-tiled launch
// (shared by dspmm_batch_* and the serving loop).
#pragma once

#include <vector>

#include "dspmm.h"
#include "internal.hpp"
#include "native.hpp"

namespace dspmm {

struct BatchSchedule {
    std::vector<dspmm_config> cfg;       // resolved nnz-tiled config per item
    std::vector<ItemDesc> items;         // carry is left null until bound
    std::vector<int64_t> carry_off;      // floats from the start of the batch's carry area
    std::vector<size_t> carry_bytes;     // per item
    int64_t carry_floats = 0;
    std::vector<int4> segs;              // per-XCD work segments
    std::vector<int> xoff;               // 9 segment offsets + 8 unit counts
    std::vector<int> split_prefix;       // count + 1
    int nunits = 0, nsplits = 0, max_blocks_x = 0, lds_ints = 1;
    int units_x[8] = {0};
};

// One item's part of a schedule: its nnz-tiled configuration, descriptor
// (operands, strides and carry left to bind), work units, cost and carry
// size.  It depends on the matrix, nv, the dense type and the operands'
// alignment class only, so a server computes it once per (matrix, nv) and
// every later admission copies it.
struct ItemFragment {
    dspmm_config cfg{};
    ItemDesc d{};
    int units = 0;
    double cost = 0;
    int lds_ints = 1;
    int nslots = 0, nsplits = 0;
};

// The fragment of matrix A with nv columns (locks A while reading it).
dspmm_status_t schedule_item(dspmm_matrix_s *A, int nv, int64_t ldb, int64_t ldc, const void *B,
                             const void *C, dspmm_dtype_t btype, ItemFragment &f);
// A schedule from one fragment per item (frags[i] for items[i], whose
// operands and strides are bound here): carry offsets, split prefix sums and
// the per-XCD segments.
dspmm_status_t assemble_schedule(const ItemFragment *const *frags, const dspmm_batch_item *items,
                                 int count, BatchSchedule &s);
// schedule_item for every item, then assemble_schedule.
dspmm_status_t build_schedule(const dspmm_batch_item *items, int count, dspmm_dtype_t btype,
                              BatchSchedule &s);
// Workgroups of the launch: blocks_per_cu > 0 -> persistent grid.
int schedule_blocks(const BatchSchedule &s, int blocks_per_cu);

}  // namespace dspmm
