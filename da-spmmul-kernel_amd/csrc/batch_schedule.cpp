// batch_schedule.cpp -- host side of a batched nnz-tiled launch that needs
// no device: a schedule assembled from per-item fragments (carry offsets,
// split prefix sums, per-XCD work segments).  api_batch.cpp and the serving
// loop upload it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "batch_schedule.hpp"
#include "common.hpp"

namespace dspmm {

dspmm_status_t assemble_schedule(const ItemFragment *const *frags, const dspmm_batch_item *items,
                                 int count, BatchSchedule &s) {
    s = BatchSchedule{};
    s.cfg.resize((size_t)count);
    s.items.resize((size_t)count);
    s.carry_off.assign((size_t)count, 0);
    s.carry_bytes.assign((size_t)count, 0);
    s.split_prefix.assign((size_t)count + 1, 0);
    std::vector<int> units((size_t)count, 0);
    std::vector<double> cost((size_t)count, 0.0);
    for (int i = 0; i < count; i++) {
        const ItemFragment &f = *frags[i];
        const dspmm_batch_item &it = items[i];
        s.cfg[(size_t)i] = f.cfg;
        ItemDesc &d = s.items[(size_t)i];
        d = f.d;
        d.B = it.B;
        d.C = it.C;
        d.carry = nullptr;
        d.ldb = it.ldb;
        d.ldc = it.ldc;
        units[(size_t)i] = f.units;
        cost[(size_t)i] = f.cost;
        s.lds_ints = std::max(s.lds_ints, f.lds_ints);
        s.carry_off[(size_t)i] = s.carry_floats;
        s.carry_bytes[(size_t)i] = sizeof(float) * (size_t)f.nslots * (size_t)d.ldcarry;
        s.carry_floats += (int64_t)f.nslots * d.ldcarry;
        s.split_prefix[(size_t)i + 1] = s.split_prefix[(size_t)i] + f.nsplits;
    }
    s.nsplits = s.split_prefix[(size_t)count];

    // XCD assignment: matrices sorted by cost; a matrix heavier than 1/32 of
    // the batch is cut into 8 contiguous unit ranges (one per XCD), every
    // other matrix goes whole to the least-loaded XCD.
    std::vector<int> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return cost[(size_t)x] > cost[(size_t)y]; });
    const double total = std::accumulate(cost.begin(), cost.end(), 0.0);
    std::vector<std::vector<int4>> segs_x(kNumXcd);
    double load[kNumXcd] = {0};
    int ucount[kNumXcd] = {0};
    for (int i : order) {
        const int u = units[(size_t)i];
        if (u == 0) continue;
        if (cost[(size_t)i] > total / 32.0 && u >= kNumXcd) {
            int begin = 0;
            for (int x = 0; x < kNumXcd; x++) {
                const int end = (int)((int64_t)u * (x + 1) / kNumXcd);
                if (end > begin) {
                    segs_x[(size_t)x].push_back(make_int4(i, begin, end - begin, ucount[x]));
                    ucount[x] += end - begin;
                    load[x] += cost[(size_t)i] * (end - begin) / u;
                }
                begin = end;
            }
        } else {
            const int x = (int)(std::min_element(load, load + kNumXcd) - load);
            segs_x[(size_t)x].push_back(make_int4(i, 0, u, ucount[x]));
            ucount[x] += u;
            load[x] += cost[(size_t)i];
        }
    }
    s.xoff.assign(2 * kNumXcd + 1, 0);
    for (int x = 0; x < kNumXcd; x++) {
        s.xoff[(size_t)x] = (int)s.segs.size();
        s.segs.insert(s.segs.end(), segs_x[(size_t)x].begin(), segs_x[(size_t)x].end());
        s.xoff[(size_t)kNumXcd + 1 + x] = ucount[x];
        s.units_x[x] = ucount[x];
        s.nunits += ucount[x];
        s.max_blocks_x = std::max(s.max_blocks_x, div_up(ucount[x], kBlock / kWave));
    }
    s.xoff[(size_t)kNumXcd] = (int)s.segs.size();
    return DSPMM_STATUS_SUCCESS;
}

}  // namespace dspmm
