// serve.cpp -- the serving loop: a device-resident CSR cache, HBM-budgeted
// admission of dense operands, and continuous batching over stream slots.
//
// One worker thread owns all device work.  Its loop: (1) reap slots whose
// completion event has fired (requests complete, their operands return to
// the arena, their matrices are unpinned); (2) if a slot is free and
// requests are queued, admit queued requests in arrival order while their
// matrix is (or can be made) resident and their operands fit the arena,
// and launch them as one batched SpMM on that slot's stream; (3) otherwise
// sleep briefly.  Several slots are in flight at once, so a batch formed
// from requests that arrived while others ran overlaps them on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "batch_schedule.hpp"
#include "common.hpp"
#include "first_fit.hpp"
#include "internal.hpp"
#include "matrix.hpp"
#include "native.hpp"

using namespace dspmm;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t round_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// Matrices with at least this many nonzeros fill the GPU alone; the batching
// server launches their requests one at a time.
constexpr int64_t kServeAloneNnz = 16 << 20;
// Up to this many admitted requests whose own AUTO kernel is row-exact (and
// whose batch plan splits no row) launch one by one with it instead of as
// one batched launch: few requests gain nothing from the batch's static
// kernel, and launch one after another on the slot's stream.
constexpr size_t kServeOwnMax = 4;

// The operand arena: one device region, its blocks handed out first-fit
// (offsets aligned to 256 B).
class Arena {
  public:
    dspmm_status_t init(size_t bytes) {
        const size_t size = round_up(bytes, 256);
        if (size == 0) return DSPMM_STATUS_SUCCESS;
        dspmm_status_t s = hip_status(hipMalloc((void **)&base_, size));
        if (s == DSPMM_STATUS_SUCCESS) ff_.reset(size);
        return s;
    }
    void destroy() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }
    bool alloc(size_t n, size_t *off) { return ff_.alloc(n, off); }
    void release(size_t off, size_t n) { ff_.release(off, n); }
    char *ptr(size_t off) const { return base_ + off; }
    size_t size() const { return ff_.size(); }
    size_t peak() const { return ff_.peak(); }

  private:
    char *base_ = nullptr;
    FirstFit ff_;
};

// Per (matrix, nv): the matrix's part of a batch schedule, and whether a
// request that finds no company may run the matrix's own AUTO kernel
// instead -- only where that kernel sums every row in nonzero order, as the
// batch's kernel does when it splits no row, so the result has the same
// bits either way.
struct Frag {
    ItemFragment f;
    dspmm_config own{};
    size_t own_ws = 0;
    bool own_ok = false;
};

struct SMat {
    int nr = 0, nc = 0, nnz = 0;
    std::vector<int> rp, ci;
    std::vector<float> v;
    dspmm_matrix_t h = nullptr;
    size_t bytes = 0;
    uint64_t last_use = 0;
    int pins = 0;
    std::map<int, Frag> frags;         // by nv; cleared when the handle goes
    size_t estimate() const {
        return 4 * ((size_t)nr + 1) + 8 * (size_t)nnz + ((size_t)nr + (size_t)nnz) + (1 << 20);
    }
};

struct Job {
    dspmm_request r;
    int status = 0;                    // dspmm_status_t of the launch
    double t_submit = 0, t_admit = 0;
    size_t offB = 0, szB = 0, offC = 0, szC = 0, offK = 0, szK = 0;
    const void *B = nullptr;
    float *C = nullptr;
    float *carry = nullptr;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Job> jobs;
    // a batch's schedule: items, segments, XCD offsets, split prefix, each
    // section 256-byte aligned, packed in the pinned stage as in the device
    // block, so one copy uploads it
    char *d_sched = nullptr;
    char *h_stage = nullptr;   // pinned
    size_t stage_bytes = 0;
};

}  // namespace

struct dspmm_server_s {
    dspmm_server_config cfg{};
    std::vector<SMat> mats;
    size_t cache_used = 0, cache_peak = 0;
    uint64_t tick = 0;
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> queue;
    std::vector<dspmm_completion> done;
    std::map<uint64_t, std::pair<size_t, size_t>> owned_c;   // id -> (offset, bytes)
    std::thread worker;
    bool stop = false;
    int inflight = 0;
    std::vector<dspmm_matrix_t> evicted;                       // freed outside the mutex
    dspmm_server_stats stats{};
    int seg_cap = 0;

    size_t elem_b() const { return cfg.btype == DSPMM_R_16BF ? 2 : 4; }

    // ---- cache ---------------------------------------------------------------
    bool ensure_resident(SMat &m) {
        if (m.h) {
            m.last_use = ++tick;
            stats.cache_hits++;
            return true;
        }
        const size_t est = m.estimate();
        while (cache_used + est > cfg.cache_bytes) {
            SMat *victim = nullptr;
            for (auto &x : mats)
                if (x.h && x.pins == 0 && (!victim || x.last_use < victim->last_use)) victim = &x;
            if (!victim) return false;
            // unpinned: no launch in flight reads it.  It is freed by the
            // worker after it lets go of the mutex (hipFree may wait for
            // the device), so submit / poll callers are never held up.
            evicted.push_back(victim->h);
            victim->h = nullptr;
            victim->frags.clear();
            cache_used -= victim->bytes;
            victim->bytes = 0;
            stats.evictions++;
        }
        if (dspmm_matrix_create(&m.h, m.nr, m.nc, m.nnz, m.rp.data(), nullptr, m.ci.data(),
                                m.v.data(), 0) != DSPMM_STATUS_SUCCESS) {
            m.h = nullptr;
            return false;
        }
        m.bytes = est;
        cache_used += m.bytes;
        cache_peak = std::max(cache_peak, cache_used);
        m.last_use = ++tick;
        stats.cache_misses++;
        return true;
    }

    // ---- admission -------------------------------------------------------------
    // Reserves arena space for a queued job; false if it does not fit now.
    // A request on a matrix that fills the GPU by itself is launched alone,
    // with the matrix's own AUTO configuration: batching cannot make it
    // faster, and the batch's static configuration makes it slower.
    bool runs_alone(const SMat &m) const {
        return cfg.plain == 0 && (int64_t)m.nnz >= (cfg.alone_nnz > 0 ? cfg.alone_nnz : kServeAloneNnz);
    }

    // The (matrix, nv) fragment, computed on the first request and kept
    // while the matrix stays resident: admission copies it.  The arena's
    // operands are 256-byte aligned with ld = nv, so one alignment class.
    Frag *fragment(SMat &m, int nv) {
        auto it = m.frags.find(nv);
        if (it != m.frags.end()) return &it->second;
        Frag fr;
        if (schedule_item(m.h, nv, nv, nv, reinterpret_cast<const void *>(256),
                          reinterpret_cast<const void *>(256), cfg.btype, fr.f) != DSPMM_STATUS_SUCCESS)
            return nullptr;
        if (dspmm_spmm_ex_resolve(m.h, DENSE_ROW_MAJOR, nv, cfg.btype, nullptr, &fr.own) ==
                DSPMM_STATUS_SUCCESS &&
            dspmm_spmm_ex_workspace_size(m.h, DENSE_ROW_MAJOR, nv, cfg.btype, &fr.own, &fr.own_ws) ==
                DSPMM_STATUS_SUCCESS)
            fr.own_ok = fr.f.nsplits == 0 && row_exact(m.h, nv, fr.own);
        // the own kernel's plan (a ring plan, say) built now, at admission,
        // not inside a launch (placeholder operands: only their alignment
        // class and strides matter)
        if (fr.own_ok &&
            dspmm_matrix_prepare(m.h, DENSE_ROW_MAJOR, nv, reinterpret_cast<const void *>(256), nv,
                                 cfg.btype, reinterpret_cast<const float *>(256), nv, &fr.own,
                                 0) != DSPMM_STATUS_SUCCESS)
            fr.own_ok = false;
        return &m.frags.emplace(nv, fr).first->second;
    }

    // Kernels that sum every row in nonzero order with a separately rounded
    // multiply and add (the nnz-tiled kernel when its plan splits no row).
    bool row_exact(dspmm_matrix_t h, int nv, const dspmm_config &c) const {
        switch (c.variant) {
        case DSPMM_VARIANT_CSR_SCALAR:
        case DSPMM_VARIANT_NNZ_CACHED:
        case DSPMM_VARIANT_ROW_BLOCKED:
        case DSPMM_VARIANT_RING:
            return true;
        case DSPMM_VARIANT_NNZ_TILED: {
            dspmm_config cc = c;
            complete_nb(cc, nv, nv, nv, reinterpret_cast<const void *>(256),
                        reinterpret_cast<const void *>(256), cfg.btype == DSPMM_R_16BF);
            std::lock_guard<std::mutex> lk(h->mu);
            NbPlan *plan = nullptr;
            return get_plan(h, cc.tile_nnz, &plan) == DSPMM_STATUS_SUCCESS && plan->nsplits == 0;
        }
        default:
            return false;
        }
    }

    bool reserve(Job &j, SMat &m) {
        const dspmm_request &r = j.r;
        j.szB = r.B ? 0 : (size_t)m.nc * (size_t)r.nv * elem_b();
        j.szC = r.C ? 0 : (size_t)m.nr * (size_t)r.nv * sizeof(float);
        size_t carry = 0;
        if (runs_alone(m)) {
            // its own AUTO configuration's workspace (slots run concurrently,
            // so the handle's internal buffer cannot be shared)
            (void)dspmm_spmm_ex_workspace_size(m.h, DENSE_ROW_MAJOR, r.nv, cfg.btype, nullptr,
                                               &carry);
        } else if (const Frag *fr = fragment(m, r.nv)) {
            carry = sizeof(float) * (size_t)fr->f.nslots * (size_t)fr->f.d.ldcarry;
            if (fr->own_ok) carry = std::max(carry, fr->own_ws);
        }
        j.szK = carry;
        bool okB = true, okC = true, okK = true;
        if (j.szB) okB = arena.alloc(j.szB, &j.offB);
        if (okB && j.szC) okC = arena.alloc(j.szC, &j.offC);
        if (okB && okC && j.szK) okK = arena.alloc(j.szK, &j.offK);
        if (okB && okC && okK) {
            j.B = r.B ? r.B : arena.ptr(j.offB);
            j.C = r.C ? r.C : reinterpret_cast<float *>(arena.ptr(j.offC));
            j.carry = j.szK ? reinterpret_cast<float *>(arena.ptr(j.offK)) : nullptr;
            return true;
        }
        if (okB && okC && j.szC) arena.release(j.offC, j.szC);
        if (okB && j.szB) arena.release(j.offB, j.szB);
        return false;
    }

    // ---- launch ------------------------------------------------------------------
    bool launch(Slot &sl) {
        std::vector<Job> jobs;
        const int cap = cfg.plain ? 1 : cfg.max_batch;
        bool alone = false;
        while (!queue.empty() && (int)jobs.size() < cap) {
            Job &j = queue.front();
            SMat &m = mats[(size_t)j.r.matrix];
            alone = runs_alone(m);
            if (alone && !jobs.empty()) {            // it starts its own launch next time
                alone = false;
                break;
            }
            if (!ensure_resident(m)) break;
            m.pins++;
            if (!reserve(j, m)) {
                m.pins--;
                break;
            }
            jobs.push_back(j);
            queue.pop_front();
            if (alone) break;
        }
        if (jobs.empty()) {
            // Nothing in flight can free cache or arena space: the head request
            // will never fit, so it completes with ADMISSION_REJECTED.
            if (inflight == 0 && !queue.empty()) {
                Job j = queue.front();
                queue.pop_front();
                dspmm_completion c{};
                c.id = j.r.id;
                c.status = DSPMM_STATUS_ADMISSION_REJECTED;
                c.t_submit = j.t_submit;
                c.t_admit = c.t_done = now_s();
                done.push_back(c);
                stats.rejected++;
                cv_done.notify_all();
                return true;
            }
            return false;
        }
        const double t_admit = now_s();
        std::vector<dspmm_batch_item> items;
        for (auto &j : jobs) {
            j.t_admit = t_admit;
            SMat &m = mats[(size_t)j.r.matrix];
            items.push_back(dspmm_batch_item{m.h, j.r.nv, j.B, j.r.nv, j.C, j.r.nv});
            if (!j.r.B)
                launch_fill_uniform(const_cast<void *>(j.B), (int64_t)m.nc * j.r.nv, j.r.seed,
                                    cfg.btype == DSPMM_R_16BF, sl.st);
        }
        BatchSchedule sch;
        std::vector<const ItemFragment *> fp;
        std::vector<const Frag *> own;                 // per job, while all may run their own kernel
        dspmm_status_t st = DSPMM_STATUS_SUCCESS;
        if (!alone && !cfg.plain) {
            for (auto &j : jobs) {
                const Frag *fr = fragment(mats[(size_t)j.r.matrix], j.r.nv);
                if (!fr) {
                    st = DSPMM_STATUS_NOT_SUPPORTED;
                    break;
                }
                fp.push_back(&fr->f);
                if (fr->own_ok && fr->own_ws <= j.szK) own.push_back(fr);
            }
            if (own.size() != jobs.size() || jobs.size() > kServeOwnMax) own.clear();
            if (st == DSPMM_STATUS_SUCCESS && own.empty())
                st = assemble_schedule(fp.data(), items.data(), (int)items.size(), sch);
        } else {
            st = build_schedule(items.data(), (int)items.size(), cfg.btype, sch);
        }
        if (st == DSPMM_STATUS_SUCCESS) {
            for (size_t i = 0; i < jobs.size() && i < sch.items.size(); i++) sch.items[i].carry = jobs[i].carry;
            if (!own.empty()) {
                // a few requests, each on a row-exact kernel of its own
                // matrix: one launch each on the slot's stream (the same bits
                // as the batched launch; at low load, lower latency)
                for (size_t i = 0; i < jobs.size() && st == DSPMM_STATUS_SUCCESS; i++) {
                    const Job &j = jobs[i];
                    st = dspmm_spmm_ex(items[i].A, DENSE_ROW_MAJOR, j.r.nv, j.B, j.r.nv, cfg.btype,
                                       j.C, j.r.nv, &own[i]->own, j.carry, j.szK,
                                       reinterpret_cast<dspmm_stream_t>(sl.st));
                }
            } else if (alone) {
                const Job &j = jobs[0];
                st = dspmm_spmm_ex(items[0].A, DENSE_ROW_MAJOR, j.r.nv, j.B, j.r.nv, cfg.btype, j.C,
                                   j.r.nv, nullptr, j.carry, j.szK,
                                   reinterpret_cast<dspmm_stream_t>(sl.st));
            } else if (cfg.plain == 2) {
                // the matrix's own AUTO configuration and workspace (requests
                // run in order on one stream, so the handle's buffer is not shared
                // by two launches at once)
                const Job &j = jobs[0];
                st = dspmm_spmm_ex(items[0].A, DENSE_ROW_MAJOR, j.r.nv, j.B, j.r.nv, cfg.btype, j.C,
                                   j.r.nv, nullptr, nullptr, 0,
                                   reinterpret_cast<dspmm_stream_t>(sl.st));
            } else if (cfg.plain) {
                const Job &j = jobs[0];
                st = dspmm_spmm_ex(items[0].A, DENSE_ROW_MAJOR, j.r.nv, j.B, j.r.nv, cfg.btype, j.C,
                                   j.r.nv, &sch.cfg[0], j.carry, j.szK,
                                   reinterpret_cast<dspmm_stream_t>(sl.st));
            } else {
                st = launch_batch(sl, sch, (int)items.size());
            }
        }
        // The event is recorded even after an error, so the jobs complete (with
        // the error) once the stream has drained.
        const dspmm_status_t se = hip_status(hipEventRecord(sl.ev, sl.st));
        if (st == DSPMM_STATUS_SUCCESS) st = se;
        for (auto &j : jobs) j.status = (int)st;
        sl.jobs.swap(jobs);
        sl.busy = true;
        inflight++;
        stats.batches++;
        return true;
    }

    dspmm_status_t launch_batch(Slot &sl, const BatchSchedule &sch, int count) {
        if ((int)sch.segs.size() > seg_cap) return DSPMM_STATUS_NOT_SUPPORTED;
        char *p = sl.h_stage;
        const size_t bi = sizeof(ItemDesc) * (size_t)count, bs = sizeof(int4) * sch.segs.size(),
                     bx = sizeof(int) * sch.xoff.size(), bp = sizeof(int) * sch.split_prefix.size();
        const size_t os = round_up(bi, 256), ox = os + round_up(bs, 256), op = ox + round_up(bx, 256);
        if (op + bp > sl.stage_bytes) return DSPMM_STATUS_NOT_SUPPORTED;
        std::memcpy(p, sch.items.data(), bi);
        std::memcpy(p + os, sch.segs.data(), bs);
        std::memcpy(p + ox, sch.xoff.data(), bx);
        std::memcpy(p + op, sch.split_prefix.data(), bp);
        // one upload per batch (a small copy costs microseconds of the
        // stream's time whatever its size)
        const dspmm_status_t st =
            hip_status(hipMemcpyAsync(sl.d_sched, p, op + bp, hipMemcpyHostToDevice, sl.st));
        if (st != DSPMM_STATUS_SUCCESS) return st;
        BatchLaunch a;
        a.items = reinterpret_cast<const ItemDesc *>(sl.d_sched);
        a.segs = reinterpret_cast<const int4 *>(sl.d_sched + os);
        a.xoff = reinterpret_cast<const int *>(sl.d_sched + ox);
        a.lds_ints_per_wave = sch.lds_ints;
        a.nblocks = schedule_blocks(sch, 0);
        a.split_prefix = reinterpret_cast<const int *>(sl.d_sched + op);
        a.count = count;
        a.nsplits = sch.nsplits;
        a.bf16 = cfg.btype == DSPMM_R_16BF;
        return launch_nnz_tiled_batched(a, sl.st);
    }

    // ---- completion ------------------------------------------------------------
    bool reap() {
        bool any = false;
        for (auto &sl : slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.ev);
            if (q == hipErrorNotReady) continue;
            const double t = now_s();
            for (auto &j : sl.jobs) {
                dspmm_completion c{};
                c.id = j.r.id;
                c.status = j.status ? j.status : (q == hipSuccess ? 0 : (int)DSPMM_STATUS_HIP_ERROR);
                c.batch_size = (int)sl.jobs.size();
                c.t_submit = j.t_submit;
                c.t_admit = j.t_admit;
                c.t_done = t;
                c.C = j.C;
                if (j.szB) arena.release(j.offB, j.szB);
                if (j.szK) arena.release(j.offK, j.szK);
                if (j.szC) {
                    if (j.r.flags & DSPMM_REQUEST_DISCARD_RESULT) {
                        arena.release(j.offC, j.szC);
                        c.C = nullptr;
                    } else {
                        owned_c[j.r.id] = {j.offC, j.szC};
                    }
                }
                done.push_back(c);
                mats[(size_t)j.r.matrix].pins--;
                stats.completed++;
            }
            sl.jobs.clear();
            sl.busy = false;
            inflight--;
            any = true;
        }
        if (any) cv_done.notify_all();
        return any;
    }

    Slot *free_slot() {
        for (auto &sl : slots)
            if (!sl.busy) return &sl;
        return nullptr;
    }

    void run() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool progressed = reap();
            if (stop && queue.empty() && inflight == 0) break;
            Slot *sl = free_slot();
            if (sl && !queue.empty() && launch(*sl)) progressed = true;
            if (!evicted.empty()) {
                std::vector<dspmm_matrix_t> gone;
                gone.swap(evicted);
                lk.unlock();
                for (dspmm_matrix_t h : gone) dspmm_matrix_destroy(h);
                lk.lock();
            }
            if (!progressed) {
                if (inflight > 0) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(5));
                    lk.lock();
                } else {
                    cv_work.wait_for(lk, std::chrono::milliseconds(2));
                }
            }
        }
    }
};

extern "C" {

double dspmm_server_now(void) { return now_s(); }

dspmm_status_t dspmm_server_create(dspmm_server_t *out, const dspmm_server_config *cfg) {
    if (!out || !cfg || cfg->max_inflight < 1 || cfg->max_batch < 1) return DSPMM_STATUS_INVALID_VALUE;
    if (cfg->btype != DSPMM_R_32F && cfg->btype != DSPMM_R_16BF) return DSPMM_STATUS_INVALID_VALUE;
    if (cfg->plain < 0 || cfg->plain > 2 || cfg->alone_nnz < 0) return DSPMM_STATUS_INVALID_VALUE;
    *out = nullptr;
    auto *s = new dspmm_server_s;
    s->cfg = *cfg;
    dspmm_status_t st = s->arena.init(cfg->operand_bytes);
    const int nslots = cfg->plain ? std::max(cfg->max_inflight, 64) : cfg->max_inflight;
    const int items_cap = cfg->plain ? 1 : cfg->max_batch;
    s->seg_cap = items_cap * kNumXcd + kNumXcd;
    s->slots.resize((size_t)nslots);
    for (size_t i = 0; i < s->slots.size() && st == DSPMM_STATUS_SUCCESS; i++) {
        Slot &sl = s->slots[i];
        if (cfg->plain && i > 0) {
            sl.st = s->slots[0].st;                    // plain: one stream, in order
        } else {
            st = hip_status(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking));
        }
        if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
        if (cfg->plain) continue;
        const size_t bi = sizeof(ItemDesc) * (size_t)items_cap, bs = sizeof(int4) * (size_t)s->seg_cap,
                     bx = sizeof(int) * (2 * kNumXcd + 1), bp = sizeof(int) * ((size_t)items_cap + 1);
        sl.stage_bytes = round_up(bi, 256) + round_up(bs, 256) + round_up(bx, 256) + bp;
        if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipMalloc((void **)&sl.d_sched, sl.stage_bytes));
        if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipHostMalloc((void **)&sl.h_stage, sl.stage_bytes));
    }
    if (st != DSPMM_STATUS_SUCCESS) {
        dspmm_server_destroy(s);
        return st;
    }
    s->worker = std::thread([s]() { s->run(); });
    *out = s;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_server_destroy(dspmm_server_t s) {
    if (!s) return DSPMM_STATUS_SUCCESS;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    (void)hipDeviceSynchronize();
    for (size_t i = 0; i < s->slots.size(); i++) {
        Slot &sl = s->slots[i];
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st && !(s->cfg.plain && i > 0)) (void)hipStreamDestroy(sl.st);
        (void)hipFree(sl.d_sched);
        if (sl.h_stage) (void)hipHostFree(sl.h_stage);
    }
    for (auto &m : s->mats)
        if (m.h) dspmm_matrix_destroy(m.h);
    for (dspmm_matrix_t h : s->evicted) dspmm_matrix_destroy(h);
    s->arena.destroy();
    delete s;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_server_add_matrix(dspmm_server_t s, int nr, int nc, int nnz,
                                       const int *h_rowPtr, const int *h_colIdx,
                                       const float *h_values, int preload, int *matrix_id) {
    if (!s || !matrix_id || nr < 0 || nc < 0 || nnz < 0 || !h_rowPtr) return DSPMM_STATUS_INVALID_VALUE;
    if (nnz > 0 && (!h_colIdx || !h_values)) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    SMat m;
    m.nr = nr;
    m.nc = nc;
    m.nnz = nnz;
    m.rp.assign(h_rowPtr, h_rowPtr + (size_t)nr + 1);
    m.ci.assign(h_colIdx, h_colIdx + nnz);
    m.v.assign(h_values, h_values + nnz);
    s->mats.push_back(std::move(m));
    *matrix_id = (int)s->mats.size() - 1;
    if (preload && !s->ensure_resident(s->mats.back())) return DSPMM_STATUS_ADMISSION_REJECTED;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_server_submit(dspmm_server_t s, const dspmm_request *req) {
    const double t = now_s();
    if (!s || !req) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    if (req->matrix < 0 || req->matrix >= (int)s->mats.size() || req->nv < 1)
        return DSPMM_STATUS_INVALID_VALUE;
    const SMat &m = s->mats[(size_t)req->matrix];
    const size_t need = (req->B ? 0 : (size_t)m.nc * req->nv * s->elem_b()) +
                        (req->C ? 0 : (size_t)m.nr * req->nv * sizeof(float));
    if (m.estimate() > s->cfg.cache_bytes || need > s->arena.size()) {
        s->stats.rejected++;
        return DSPMM_STATUS_ADMISSION_REJECTED;
    }
    Job j;
    j.r = *req;
    j.t_submit = t;
    s->queue.push_back(j);
    s->stats.submitted++;
    s->cv_work.notify_one();
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_server_drain(dspmm_server_t s) {
    if (!s) return DSPMM_STATUS_INVALID_VALUE;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [s]() { return s->queue.empty() && s->inflight == 0; });
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_server_replay(dspmm_server_t s, const dspmm_request *req, const double *at,
                                   int n) {
    if (!s || n < 0 || (n > 0 && (!req || !at))) return DSPMM_STATUS_INVALID_VALUE;
    const double t0 = now_s();
    for (int i = 0; i < n; i++) {
        const double target = t0 + at[i];
        for (;;) {
            const double d = target - now_s();
            if (d <= 0) break;
            if (d > 300e-6) std::this_thread::sleep_for(std::chrono::duration<double>(d - 200e-6));
        }
        const dspmm_status_t st = dspmm_server_submit(s, &req[i]);
        if (st != DSPMM_STATUS_SUCCESS && st != DSPMM_STATUS_ADMISSION_REJECTED) return st;
    }
    return dspmm_server_drain(s);
}

int dspmm_server_poll(dspmm_server_t s, dspmm_completion *out, int max) {
    if (!s || !out || max <= 0) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    const int n = std::min<int>(max, (int)s->done.size());
    std::copy(s->done.begin(), s->done.begin() + n, out);
    s->done.erase(s->done.begin(), s->done.begin() + n);
    return n;
}

dspmm_status_t dspmm_server_release(dspmm_server_t s, uint64_t id) {
    if (!s) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->owned_c.find(id);
    if (it == s->owned_c.end()) return DSPMM_STATUS_INVALID_VALUE;
    s->arena.release(it->second.first, it->second.second);
    s->owned_c.erase(it);
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_server_get_stats(dspmm_server_t s, dspmm_server_stats *st) {
    if (!s || !st) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    st->cache_used = s->cache_used;
    st->cache_peak = s->cache_peak;
    st->operand_peak = s->arena.peak();
    st->mean_batch = s->stats.batches ? (double)s->stats.completed / (double)s->stats.batches : 0.0;
    return DSPMM_STATUS_SUCCESS;
}

}  // extern "C"
