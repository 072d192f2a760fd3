// api_batch.cpp -- C-ABI of the batched path: one persistent launch over
// the nnz tiles of many matrices, its hipGraph replay, and the one-launch-
// per-matrix baseline it is compared against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "matrix.hpp"
#include "native.hpp"

using namespace dspmm;

struct dspmm_batch_s {
    int count = 0;
    dspmm_dtype_t btype = DSPMM_R_32F;
    std::vector<dspmm_batch_item> items;
    std::vector<dspmm_config> cfg;        // resolved per-item configs
    std::vector<ItemDesc> h_items;
    std::vector<int64_t> carry_off;       // per item, floats
    std::vector<size_t> carry_bytes_item;
    ItemDesc *d_items = nullptr;
    int4 *d_segs = nullptr;
    int *d_xoff = nullptr;
    int4 *d_splits = nullptr;
    float *carry = nullptr;
    size_t carry_bytes = 0;
    size_t table_bytes = 0;
    int nsegs = 0, nsplits = 0, nunits = 0, nblocks = 0, lds_ints = 0;
    int units_x[kNumXcd] = {0};
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
    (void)hipFree(b->d_splits);
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

int device_cus() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus > 0 ? cus : 256;
}

dspmm_status_t run_batched(dspmm_batch_s *b, hipStream_t st, hipEvent_t mid) {
    BatchLaunch a;
    a.items = b->d_items;
    a.segs = b->d_segs;
    a.xoff = b->d_xoff;
    a.lds_ints_per_wave = b->lds_ints;
    a.nblocks = b->nunits > 0 ? b->nblocks : 0;
    a.splits = b->d_splits;
    a.nsplits = b->nsplits;
    a.bf16 = b->btype == DSPMM_R_16BF;
    a.mid_event = mid;
    return launch_nnz_tiled_batched(a, st);
}

dspmm_status_t run_plain(dspmm_batch_s *b, hipStream_t st) {
    for (int i = 0; i < b->count; i++) {
        const dspmm_batch_item &it = b->items[(size_t)i];
        float *ws = b->carry ? b->carry + b->carry_off[(size_t)i] : nullptr;
        dspmm_status_t s = dspmm_spmm_ex(it.A, DENSE_ROW_MAJOR, it.nv, it.B, it.ldb, b->btype, it.C,
                                         it.ldc, &b->cfg[(size_t)i], ws,
                                         b->carry_bytes_item[(size_t)i],
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
    const bool bf16 = btype == DSPMM_R_16BF;
    auto *b = new dspmm_batch_s;
    b->count = count;
    b->btype = btype;
    b->items.assign(items, items + count);
    b->cfg.resize((size_t)count);
    b->h_items.resize((size_t)count);
    b->carry_off.assign((size_t)count, 0);
    b->carry_bytes_item.assign((size_t)count, 0);

    std::vector<int> units((size_t)count, 0);
    std::vector<double> cost((size_t)count, 0.0);
    std::vector<int4> splits;
    int64_t carry_floats = 0;
    int max_lds = 1;
    for (int i = 0; i < count; i++) {
        const dspmm_batch_item &it = items[i];
        dspmm_matrix_s *A = it.A;
        if (!A || it.nv < 1 || it.ldb < it.nv || it.ldc < it.nv || (A->nr > 0 && !it.C) ||
            (A->nnz > 0 && !it.B)) {
            free_batch(b);
            return DSPMM_STATUS_INVALID_VALUE;
        }
        std::lock_guard<std::mutex> lk(A->mu);
        dspmm_config c = resolve(A, DENSE_ROW_MAJOR, it.nv, btype, nullptr);
        if (c.variant != DSPMM_VARIANT_NNZ_TILED) {
            c = dspmm_config{};
            c.variant = DSPMM_VARIANT_NNZ_TILED;
        }
        complete_nb(c, it.nv, it.ldb, it.ldc, it.B, it.C, bf16);
        if (nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) > kMaxTileLds) {
            c.tile_nnz = 16;
            if (nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) > kMaxTileLds) {
                free_batch(b);
                return DSPMM_STATUS_NOT_SUPPORTED;
            }
        }
        NbPlan *plan = nullptr;
        dspmm_status_t s = get_plan(A, c.tile_nnz, &plan);
        if (s != DSPMM_STATUS_SUCCESS) {
            free_batch(b);
            return s;
        }
        b->cfg[(size_t)i] = c;
        ItemDesc &d = b->h_items[(size_t)i];
        d.tiles = plan->d_tiles;
        d.rowPtr = A->rowPtr;
        d.colIdx = A->colIdx;
        d.values = A->values;
        d.B = it.B;
        d.C = it.C;
        d.carry = nullptr;                          // patched after allocation
        d.ldb = it.ldb;
        d.ldc = it.ldc;
        d.ldcarry = carry_ld(it.nv);
        d.nv = it.nv;
        d.slab_w = c.vec * c.lanes;
        d.S = c.tile_nnz;
        d.cls = item_class(c.vec, c.lanes);
        d.ntiles = plan->ntiles;
        d.nslabs = div_up(it.nv, d.slab_w);
        const int G = kWave / c.lanes;
        units[(size_t)i] = d.nslabs * div_up(plan->ntiles, G);
        cost[(size_t)i] = (double)A->nnz * (8.0 + (double)it.nv * (bf16 ? 2 : 4)) +
                          (double)A->nr * it.nv * 4.0;
        max_lds = std::max(max_lds, (int)(nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) / (kBlock / kWave) / 4));
        b->carry_off[(size_t)i] = carry_floats;
        b->carry_bytes_item[(size_t)i] = sizeof(float) * (size_t)plan->nslots * (size_t)d.ldcarry;
        carry_floats += (int64_t)plan->nslots * d.ldcarry;
        for (const int4 &sp : plan->h_splits) splits.push_back(make_int4(i, sp.y, sp.z, sp.w));
    }
    b->lds_ints = max_lds;
    b->nsplits = (int)splits.size();

    // XCD assignment: matrices sorted by cost; a matrix heavier than 1/32 of
    // the batch is cut into 8 contiguous unit ranges (one per XCD), every
    // other matrix goes whole to the least-loaded XCD.
    std::vector<int> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[(size_t)x] > cost[(size_t)y]; });
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
            int x = (int)(std::min_element(load, load + kNumXcd) - load);
            segs_x[(size_t)x].push_back(make_int4(i, 0, u, ucount[x]));
            ucount[x] += u;
            load[x] += cost[(size_t)i];
        }
    }
    std::vector<int4> segs;
    std::vector<int> xoff(2 * kNumXcd + 1, 0);
    int max_blocks_x = 0;
    for (int x = 0; x < kNumXcd; x++) {
        xoff[(size_t)x] = (int)segs.size();
        segs.insert(segs.end(), segs_x[(size_t)x].begin(), segs_x[(size_t)x].end());
        xoff[(size_t)kNumXcd + 1 + x] = ucount[x];
        b->units_x[x] = ucount[x];
        b->nunits += ucount[x];
        max_blocks_x = std::max(max_blocks_x, div_up(ucount[x], kBlock / kWave));
    }
    xoff[(size_t)kNumXcd] = (int)segs.size();
    b->nsegs = (int)segs.size();

    const size_t block_lds = (size_t)(kBlock / kWave) * (size_t)b->lds_ints * sizeof(int);
    int bpc = blocks_per_cu > 0 ? blocks_per_cu : 3;
    bpc = std::max(1, std::min<int>(bpc, (int)((160 * 1024) / std::max<size_t>(block_lds, 1))));
    const int per_x = std::max(1, device_cus() / kNumXcd) * bpc;
    b->nblocks = kNumXcd * std::min(per_x, std::max(1, max_blocks_x));

    dspmm_status_t st = DSPMM_STATUS_SUCCESS;
    if (carry_floats > 0) {
        b->carry_bytes = sizeof(float) * (size_t)carry_floats;
        st = hip_status(hipMalloc((void **)&b->carry, b->carry_bytes));
    }
    for (int i = 0; i < count && st == DSPMM_STATUS_SUCCESS; i++)
        b->h_items[(size_t)i].carry = b->carry ? b->carry + b->carry_off[(size_t)i] : nullptr;
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_items, b->h_items);
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_segs, segs);
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_xoff, xoff);
    if (st == DSPMM_STATUS_SUCCESS) st = to_device(&b->d_splits, splits);
    if (st != DSPMM_STATUS_SUCCESS) {
        free_batch(b);
        return st;
    }
    b->table_bytes = sizeof(ItemDesc) * b->h_items.size() + sizeof(int4) * (segs.size() + splits.size()) +
                     sizeof(int) * xoff.size();
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
    info->units = b->nunits;
    info->segments = b->nsegs;
    info->splits = b->nsplits;
    info->blocks = b->nblocks;
    for (int x = 0; x < kNumXcd; x++) info->units_per_xcd[x] = b->units_x[x];
    info->table_bytes = b->table_bytes;
    info->carry_bytes = b->carry_bytes;
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
    const dspmm_config &c = b->cfg[(size_t)index];
    const int eb = b->btype == DSPMM_R_16BF ? 2 : 4;
    if (!aligned(B, c.vec * eb) || !aligned(C, c.vec * 4)) return DSPMM_STATUS_NOT_SUPPORTED;
    it.B = B;
    it.C = C;
    ItemDesc &d = b->h_items[(size_t)index];
    d.B = B;
    d.C = C;
    // Ordered after earlier launches on the stream that read the old entry.
    return hip_status(hipMemcpyAsync(b->d_items + index, &d, sizeof(ItemDesc),
                                     hipMemcpyHostToDevice, as_hip(stream)));
}

dspmm_status_t dspmm_benchmark_batch(dspmm_batch_t b, int mode, int warmup, int reps,
                                     float *ms_mean, float *ms_per_launch, dspmm_stream_t stream) {
    if (!b || reps < 1 || warmup < 0 || mode < 0 || mode > 2) return DSPMM_STATUS_INVALID_VALUE;
    const hipStream_t hs = as_hip(stream);
    auto once = [&]() -> dspmm_status_t {
        if (mode == 0) return run_batched(b, hs, nullptr);
        if (mode == 1) return dspmm_batch_graph_launch(b, stream);
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
