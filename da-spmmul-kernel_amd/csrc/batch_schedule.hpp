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

// Resolves each item's config (its matrix's autotuned or heuristic
// nnz-tiled config), fetches the tile plans and cuts the work into
// per-XCD segments.  Locks each matrix while reading it.
dspmm_status_t build_schedule(const dspmm_batch_item *items, int count, dspmm_dtype_t btype,
                              BatchSchedule &s);
// Workgroups of the launch: blocks_per_cu > 0 -> persistent grid.
int schedule_blocks(const BatchSchedule &s, int blocks_per_cu);

}  // namespace dspmm
