// api_batch.cpp -- C-ABI of the batched path: one launch over the nnz tiles
// of many matrices, its hipGraph replay, and the one-launch-per-matrix
// baseline it is compared against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "batch_schedule.hpp"
#include "common.hpp"
#include "internal.hpp"
#include "matrix.hpp"
#include "native.hpp"

namespace dspmm {

dspmm_status_t schedule_item(dspmm_matrix_s *A, int nv, int64_t ldb, int64_t ldc, const void *B,
                             const void *C, dspmm_dtype_t btype, ItemFragment &f) {
    const bool bf16 = btype == DSPMM_R_16BF;
    f = ItemFragment{};
    if (!A || nv < 1 || ldb < nv || ldc < nv) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(A->mu);
    // The static nnz-tiled default: configurations tuned on single
    // launches (where one matrix has the whole GPU and its caches) make
    // the batch slower, not faster.
    dspmm_config c{};
    c.variant = DSPMM_VARIANT_NNZ_TILED;
    complete_nb(c, nv, ldb, ldc, B, C, bf16);
    if (nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) > kMaxTileLds) {
        c.tile_nnz = 16;
        if (nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) > kMaxTileLds) return DSPMM_STATUS_NOT_SUPPORTED;
    }
    NbPlan *plan = nullptr;
    dspmm_status_t st = get_plan(A, c.tile_nnz, &plan);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    f.cfg = c;
    ItemDesc &d = f.d;
    d.tiles = plan->d_tiles;
    d.rowPtr = A->rowPtr;
    d.colIdx = A->colIdx;
    d.values = A->values;
    d.ldcarry = carry_ld(nv);
    d.nv = nv;
    d.slab_w = c.vec * c.lanes;
    d.S = c.tile_nnz;
    d.cls = item_class(c.vec, c.lanes);
    d.ntiles = plan->ntiles;
    d.nslabs = div_up(nv, d.slab_w);
    d.splits = plan->d_splits;
    d.nsplits = plan->nsplits;
    const int G = kWave / c.lanes;
    f.units = d.nslabs * div_up(plan->ntiles, G);
    f.cost = (double)A->nnz * (8.0 + (double)nv * (bf16 ? 2 : 4)) + (double)A->nr * nv * 4.0;
    f.lds_ints = (int)(nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) / (kBlock / kWave) / 4);
    f.nslots = plan->nslots;
    f.nsplits = plan->nsplits;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t build_schedule(const dspmm_batch_item *items, int count, dspmm_dtype_t btype,
                              BatchSchedule &s) {
    std::vector<ItemFragment> frags((size_t)count);
    std::vector<const ItemFragment *> fp((size_t)count);
    for (int i = 0; i < count; i++) {
        const dspmm_batch_item &it = items[i];
        if (!it.A || (it.A->nr > 0 && !it.C) || (it.A->nnz > 0 && !it.B)) return DSPMM_STATUS_INVALID_VALUE;
        dspmm_status_t st = schedule_item(it.A, it.nv, it.ldb, it.ldc, it.B, it.C, btype, frags[(size_t)i]);
        if (st != DSPMM_STATUS_SUCCESS) return st;
        fp[(size_t)i] = &frags[(size_t)i];
    }
    return assemble_schedule(fp.data(), items, count, s);
}

int schedule_blocks(const BatchSchedule &s, int blocks_per_cu) {
    if (s.nunits == 0) return 0;
    if (blocks_per_cu <= 0) return kNumXcd * std::max(1, s.max_blocks_x);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t block_lds = (size_t)(kBlock / kWave) * (size_t)s.lds_ints * sizeof(int);
    const int bpc = std::max(1, std::min<int>(blocks_per_cu,
                                              (int)((160 * 1024) / std::max<size_t>(block_lds, 1))));
    const int per_x = std::max(1, cus / kNumXcd) * bpc;
    return kNumXcd * std::min(per_x, std::max(1, s.max_blocks_x));
}

}  // namespace dspmm

using namespace dspmm;

struct dspmm_batch_s {
    int count = 0;
    dspmm_dtype_t btype = DSPMM_R_32F;
    std::vector<dspmm_batch_item> items;
    BatchSchedule sch;
    ItemDesc *d_items = nullptr;
    int4 *d_segs = nullptr;
    int *d_xoff = nullptr;
    int *d_prefix = nullptr;
    float *carry = nullptr;
    size_t carry_bytes = 0;
    size_t table_bytes = 0;
    int nblocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void free_batch(dspmm_batch_s *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_segs);
    (void)hipFree(b->d_xoff);
    (void)hipFree(b->d_prefix);
    (void)hipFree(b->carry);
    delete b;
}

template <typename T>
dspmm_status_t to_device(T **dst, const std::vector<T> &src) {
    *dst = nullptr;
    if (src.empty()) return DSPMM_STATUS_SUCCESS;
    dspmm_status_t s = hip_status(hipMalloc((void **)dst, sizeof(T) * src.size()));
    if (s != DSPMM_STATUS_SUCCESS) return s;
    return hip_status(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
}

dspmm_status_t run_batched(dspmm_batch_s *b, hipStream_t st, hipEvent_t mid) {
    BatchLaunch a;
    a.items = b->d_items;
    a.segs = b->d_segs;
    a.xoff = b->d_xoff;
    a.lds_ints_per_wave = b->sch.lds_ints;
    a.nblocks = b->nblocks;
    a.split_prefix = b->d_prefix;
    a.count = b->count;
    a.nsplits = b->sch.nsplits;
    a.bf16 = b->btype == DSPMM_R_16BF;
    a.mid_event = mid;
    return launch_nnz_tiled_batched(a, st);
}

// One launch per item with the item's own AUTO configuration (its tuned
// configuration when cached, else the selector's, else the static default)
// and the matrix's own workspace.
dspmm_status_t run_plain_own(dspmm_batch_s *b, hipStream_t st) {
    for (int i = 0; i < b->count; i++) {
        const dspmm_batch_item &it = b->items[(size_t)i];
        dspmm_status_t s = dspmm_spmm_ex(it.A, DENSE_ROW_MAJOR, it.nv, it.B, it.ldb, b->btype, it.C,
                                         it.ldc, nullptr, nullptr, 0,
                                         reinterpret_cast<dspmm_stream_t>(st));
        if (s != DSPMM_STATUS_SUCCESS) return s;
    }
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t run_plain(dspmm_batch_s *b, hipStream_t st) {
    for (int i = 0; i < b->count; i++) {
        const dspmm_batch_item &it = b->items[(size_t)i];
        float *ws = b->carry ? b->carry + b->sch.carry_off[(size_t)i] : nullptr;
        dspmm_status_t s = dspmm_spmm_ex(it.A, DENSE_ROW_MAJOR, it.nv, it.B, it.ldb, b->btype, it.C,
                                         it.ldc, &b->sch.cfg[(size_t)i], ws,
                                         b->sch.carry_bytes[(size_t)i],
                                         reinterpret_cast<dspmm_stream_t>(st));
        if (s != DSPMM_STATUS_SUCCESS) return s;
    }
    return DSPMM_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

dspmm_status_t dspmm_batch_create(dspmm_batch_t *out, const dspmm_batch_item *items, int count,
                                  dspmm_dtype_t btype, int blocks_per_cu) {
    if (!out || count < 0 || (count > 0 && !items)) return DSPMM_STATUS_INVALID_VALUE;
    if (btype != DSPMM_R_32F && btype != DSPMM_R_16BF) return DSPMM_STATUS_INVALID_VALUE;
    *out = nullptr;
    auto *b = new dspmm_batch_s;
    b->count = count;
    b->btype = btype;
    b->items.assign(items, items + count);
    dspmm_status_t st = build_schedule(items, count, btype, b->sch);
    if (st != DSPMM_STATUS_SUCCESS) {
        free_batch(b);
        return st;
    }
    b->nblocks = schedule_blocks(b->sch, blocks_per_cu);
    if (b->sch.carry_floats > 0) {
        b->carry_bytes = sizeof(float) * (size_t)b->sch.carry_floats;
        st = hip_status(hipMalloc((void **)&b->carry, b->carry_bytes));
    }
    for (int i = 0; i < count && st == DSPMM_STATUS_SUCCESS; i++)
        b->sch.items[(size_t)i].carry = b->carry ? b->carry + b->sch.carry_off[(size_t)i] : nullptr;
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_items, b->sch.items);
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_segs, b->sch.segs);
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_xoff, b->sch.xoff);
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_prefix, b->sch.split_prefix);
    if (st != DSPMM_STATUS_SUCCESS) {
        free_batch(b);
        return st;
    }
    b->table_bytes = sizeof(ItemDesc) * b->sch.items.size() + sizeof(int4) * b->sch.segs.size() +
                     sizeof(int) * (b->sch.xoff.size() + b->sch.split_prefix.size());
    *out = b;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_batch_destroy(dspmm_batch_t b) {
    if (!b) return DSPMM_STATUS_SUCCESS;
    (void)hipDeviceSynchronize();   // no launch may still read the tables
    free_batch(b);
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_batch_describe(dspmm_batch_t b, dspmm_batch_info *info) {
    if (!b || !info) return DSPMM_STATUS_INVALID_VALUE;
    info->count = b->count;
    info->units = b->sch.nunits;
    info->segments = (int)b->sch.segs.size();
    info->splits = b->sch.nsplits;
    info->blocks = b->nblocks;
    for (int x = 0; x < kNumXcd; x++) info->units_per_xcd[x] = b->sch.units_x[x];
    info->table_bytes = b->table_bytes;
    info->carry_bytes = b->carry_bytes;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_batch_item_config(dspmm_batch_t b, int index, dspmm_config *out) {
    if (!b || !out || index < 0 || index >= b->count) return DSPMM_STATUS_INVALID_VALUE;
    *out = b->sch.cfg[(size_t)index];
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_batch_run(dspmm_batch_t b, dspmm_stream_t stream) {
    if (!b) return DSPMM_STATUS_INVALID_VALUE;
    return run_batched(b, as_hip(stream), nullptr);
}

dspmm_status_t dspmm_batch_graph_launch(dspmm_batch_t b, dspmm_stream_t stream) {
    if (!b) return DSPMM_STATUS_INVALID_VALUE;
    if (!b->exec) {
        hipStream_t cap = nullptr;
        dspmm_status_t st = hip_status(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        if (st != DSPMM_STATUS_SUCCESS) return st;
        st = hip_status(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        if (st == DSPMM_STATUS_SUCCESS) {
            dspmm_status_t rs = run_batched(b, cap, nullptr);
            hipGraph_t g = nullptr;
            st = hip_status(hipStreamEndCapture(cap, &g));
            if (rs != DSPMM_STATUS_SUCCESS) st = rs;
            if (st == DSPMM_STATUS_SUCCESS) {
                b->graph = g;
                st = hip_status(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        (void)hipStreamDestroy(cap);
        if (st != DSPMM_STATUS_SUCCESS) return st;
    }
    return hip_status(hipGraphLaunch(b->exec, as_hip(stream)));
}

dspmm_status_t dspmm_batch_run_plain(dspmm_batch_t b, dspmm_stream_t stream) {
    if (!b) return DSPMM_STATUS_INVALID_VALUE;
    return run_plain(b, as_hip(stream));
}

dspmm_status_t dspmm_batch_set_operands(dspmm_batch_t b, int index, const void *B, float *C,
                                        dspmm_stream_t stream) {
    if (!b || index < 0 || index >= b->count) return DSPMM_STATUS_INVALID_VALUE;
    dspmm_batch_item &it = b->items[(size_t)index];
    if ((it.A->nnz > 0 && !B) || (it.A->nr > 0 && !C)) return DSPMM_STATUS_INVALID_VALUE;
    const dspmm_config &c = b->sch.cfg[(size_t)index];
    const int eb = b->btype == DSPMM_R_16BF ? 2 : 4;
    if (!aligned(B, c.vec * eb) || !aligned(C, c.vec * 4)) return DSPMM_STATUS_NOT_SUPPORTED;
    it.B = B;
    it.C = C;
    ItemDesc &d = b->sch.items[(size_t)index];
    d.B = B;
    d.C = C;
    // Ordered after earlier launches on the stream that read the old entry.
    return hip_status(hipMemcpyAsync(b->d_items + index, &d, sizeof(ItemDesc),
                                     hipMemcpyHostToDevice, as_hip(stream)));
}

dspmm_status_t dspmm_benchmark_batch(dspmm_batch_t b, int mode, int warmup, int reps,
                                     float *ms_mean, float *ms_per_launch, dspmm_stream_t stream) {
    if (!b || reps < 1 || warmup < 0 || mode < 0 || mode > 3) return DSPMM_STATUS_INVALID_VALUE;
    const hipStream_t hs = as_hip(stream);
    auto once = [&]() -> dspmm_status_t {
        if (mode == 0) return run_batched(b, hs, nullptr);
        if (mode == 1) return dspmm_batch_graph_launch(b, stream);
        if (mode == 3) return run_plain_own(b, hs);
        return run_plain(b, hs);
    };
    dspmm_status_t st = DSPMM_STATUS_SUCCESS;
    for (int i = 0; i < warmup && st == DSPMM_STATUS_SUCCESS; i++) st = once();
    if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipStreamSynchronize(hs));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    std::vector<hipEvent_t> ev((size_t)reps + 1, nullptr);
    for (auto &e : ev) if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventCreate(&e));
    if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventRecord(ev[0], hs));
    for (int i = 0; i < reps && st == DSPMM_STATUS_SUCCESS; i++) {
        st = once();
        if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventRecord(ev[(size_t)i + 1], hs));
    }
    if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventSynchronize(ev[(size_t)reps]));
    double sum = 0;
    for (int i = 0; i < reps && st == DSPMM_STATUS_SUCCESS; i++) {
        float t = 0;
        st = hip_status(hipEventElapsedTime(&t, ev[(size_t)i], ev[(size_t)i + 1]));
        sum += t;
        if (ms_per_launch) ms_per_launch[i] = t;
    }
    if (st == DSPMM_STATUS_SUCCESS && ms_mean) *ms_mean = (float)(sum / reps);
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    return st;
}

}  // extern "C"
