// api_native.cpp -- C-ABI of the native path: device-resident CSR handle
// (the upload), the plans' device copies (built by tile_plans.cpp and
// ring_plan.cpp), dspmm_spmm_ex dispatch, the per-matrix autotuner and the
// timed benchmark loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.hpp"
#include "matrix.hpp"
#include "native.hpp"
#include "ring.hpp"

using namespace dspmm;

namespace dspmm {

void free_ring_plan(RingPlan &p) {
    (void)hipFree(p.d_tiles);
    (void)hipFree(p.d_streams);
    (void)hipFree(p.d_rows);
    (void)hipFree(p.d_fill_off);
    (void)hipFree(p.d_fills);
    (void)hipFree(p.d_pslot);
    (void)hipFree(p.d_pval);
    p = RingPlan{};
}

void free_matrix(dspmm_matrix_s *A) {
    if (!A) return;
    for (auto &kv : A->ring_plans) free_ring_plan(kv.second);
    hipFree(A->rowPtr);
    hipFree(A->rowIdx);
    hipFree(A->colIdx);
    hipFree(A->values);
    for (auto &kv : A->plans) {
        hipFree(kv.second.d_tiles);
        hipFree(kv.second.d_splits);
        hipFree(kv.second.d_ucols);
        hipFree(kv.second.d_lidx);
        hipFree(kv.second.d_boff);
        hipFree(kv.second.d_list);
        hipFree(kv.second.d_vals);
    }
    hipFree(A->carry);
    hipFree(A->tbuf);
    delete A;
}

void compute_stats(dspmm_matrix_s *A) {
    const auto &rp = A->h_rowPtr;
    A->max_len = 0;
    double s = 0, s2 = 0;
    for (int r = 0; r < A->nr; r++) {
        const int len = rp[(size_t)r + 1] - rp[(size_t)r];
        A->max_len = std::max(A->max_len, len);
        s += len;
        s2 += (double)len * len;
    }
    A->mean = A->nr ? s / A->nr : 0.0;
    const double var = A->nr ? std::max(0.0, s2 / A->nr - A->mean * A->mean) : 0.0;
    A->cv = A->mean > 0 ? std::sqrt(var) / A->mean : 0.0;
}

bool rowptr_ok(const std::vector<int> &rp, int nr, int nnz) {
    if (rp.size() != (size_t)nr + 1 || rp[0] != 0 || rp[(size_t)nr] != nnz) return false;
    for (int r = 0; r < nr; r++)
        if (rp[(size_t)r + 1] < rp[(size_t)r]) return false;
    return true;
}

// Row offsets from a sorted COO row array; same result as
// generateCsrRowPtrArray (spmvspmm_benchmark.cu:41-76) for valid input,
// and also defined for nnz == 0.
bool rowptr_from_rows(const int *rows, int nr, int nnz, std::vector<int> &rp) {
    rp.assign((size_t)nr + 1, 0);
    for (int i = 0; i < nnz; i++) {
        if (rows[i] < 0 || rows[i] >= nr) return false;
        if (i && rows[i] < rows[i - 1]) return false;
        rp[(size_t)rows[i] + 1]++;
    }
    for (int r = 0; r < nr; r++) rp[(size_t)r + 1] += rp[(size_t)r];
    return true;
}

template <typename T>
dspmm_status_t upload(T **dst, const T *src, size_t n, hipMemcpyKind kind, hipStream_t st) {
    *dst = nullptr;
    if (n == 0) return DSPMM_STATUS_SUCCESS;
    dspmm_status_t s = hip_status(hipMalloc((void **)dst, n * sizeof(T)));
    if (s != DSPMM_STATUS_SUCCESS) return s;
    if (kind == hipMemcpyHostToDevice) return hip_status(hipMemcpy(*dst, src, n * sizeof(T), kind));
    return hip_status(hipMemcpyAsync(*dst, src, n * sizeof(T), kind, st));
}

dspmm_status_t ensure_rowidx(dspmm_matrix_s *A) {
    if (A->rowIdx || A->nnz == 0) return DSPMM_STATUS_SUCCESS;
    std::vector<int> rows((size_t)A->nnz);
    for (int r = 0; r < A->nr; r++)
        for (int q = A->h_rowPtr[(size_t)r]; q < A->h_rowPtr[(size_t)r + 1]; q++) rows[(size_t)q] = r;
    return upload(&A->rowIdx, rows.data(), rows.size(), hipMemcpyHostToDevice, nullptr);
}


int cached_smax(int umax) { return std::min(4096, 4 * umax); }

double column_reuse(dspmm_matrix_s *A) {
    if (A->reuse16 >= 0) return A->reuse16;
    A->reuse16 = 1.0;
    if (A->nnz == 0 || A->nr == 0) return A->reuse16;
    // up to 256 windows of 16 consecutive rows, spread over the matrix
    const int win = 16, nwin = std::min(256, std::max(1, A->nr / win));
    std::vector<int> cols;
    int64_t nz = 0, distinct = 0;
    for (int w = 0; w < nwin; w++) {
        const int r0 = (int)((int64_t)w * (A->nr - std::min(win, A->nr)) / std::max(1, nwin - 1));
        const int r1 = std::min(A->nr, r0 + win);
        const int p0 = A->h_rowPtr[(size_t)r0], p1 = A->h_rowPtr[(size_t)r1];
        if (p1 <= p0) continue;
        cols.assign(A->h_colIdx.begin() + p0, A->h_colIdx.begin() + p1);
        std::sort(cols.begin(), cols.end());
        nz += (int64_t)cols.size();
        distinct += (int64_t)(std::unique(cols.begin(), cols.end()) - cols.begin());
    }
    if (distinct > 0) A->reuse16 = (double)nz / (double)distinct;
    return A->reuse16;
}

int shared_row_stride(dspmm_matrix_s *A) {
    if (A->rb_stride >= 1) return A->rb_stride;
    A->rb_stride = 1;
    if (A->nnz == 0 || A->nr < 3) return A->rb_stride;
    const std::vector<int> &ci = A->h_colIdx;
    const std::vector<int> &rp = A->h_rowPtr;
    const int nsamp = std::min(256, A->nr);
    std::vector<int> sample((size_t)nsamp);
    for (int k = 0; k < nsamp; k++) sample[(size_t)k] = (int)((int64_t)k * (A->nr - 1) / std::max(1, nsamp - 1));
    // candidates: the most frequent column-minus-row offsets >= 2
    std::vector<int64_t> offs;
    for (int r : sample)
        for (int p = rp[(size_t)r]; p < rp[(size_t)r + 1]; p++) {
            const int64_t d = (int64_t)ci[(size_t)p] - r;
            if (d >= 2 && d < A->nr) offs.push_back(d);
        }
    std::sort(offs.begin(), offs.end());
    std::vector<std::pair<int, int64_t>> freq;             // (-count, offset)
    for (size_t i = 0; i < offs.size();) {
        size_t j = i;
        while (j < offs.size() && offs[j] == offs[i]) j++;
        freq.emplace_back(-(int)(j - i), offs[i]);
        i = j;
    }
    std::sort(freq.begin(), freq.end());
    if (freq.size() > 16) freq.resize(16);
    // score: columns a sampled row shares with the row that far below it
    double best = 0;
    int64_t best_d = 1;
    std::vector<std::pair<int64_t, double>> scored;
    for (const auto &f : freq) {
        const int64_t d = f.second;
        int64_t shared = 0, rows = 0;
        for (int r : sample) {
            if (r + d >= A->nr) continue;
            rows++;
            int p = rp[(size_t)r], q = rp[(size_t)(r + d)];
            const int pe = rp[(size_t)r + 1], qe = rp[(size_t)(r + d) + 1];
            while (p < pe && q < qe) {
                if (ci[(size_t)p] < ci[(size_t)q]) p++;
                else if (ci[(size_t)p] > ci[(size_t)q]) q++;
                else { shared++; p++; q++; }
            }
        }
        const double score = rows ? (double)shared / (double)rows : 0.0;
        scored.emplace_back(d, score);
        if (score > best || (score == best && score > 0 && d > best_d)) {
            best = score;
            best_d = d;
        }
    }
    const double floor_score = std::max(1.0, 0.25 * (double)A->nnz / (double)A->nr);
    if (best >= floor_score) A->rb_stride = (int)best_d;
    // the best distance below it that divides it
    A->rb_stride2 = 1;
    double best2 = 0;
    for (const auto &sc : scored) {
        const int64_t d = sc.first;
        if (A->rb_stride <= 1 || d >= A->rb_stride || A->rb_stride % d || sc.second < floor_score)
            continue;
        if (sc.second > best2 || (sc.second == best2 && d > A->rb_stride2)) {
            best2 = sc.second;
            A->rb_stride2 = (int)d;
        }
    }
    return A->rb_stride;
}

int shared_row_stride2(dspmm_matrix_s *A) {
    if (A->rb_stride < 1) (void)shared_row_stride(A);
    return std::max(1, A->rb_stride2);
}

RingShape ring_shape(dspmm_matrix_s *A, int w, int h, int zchunk) {
    RingShape sh;
    sh.w = w;
    sh.h = h;
    sh.zchunk = zchunk;
    sh.s_big = shared_row_stride(A);
    sh.s_mid = shared_row_stride2(A);
    if (sh.s_big > 1 && (sh.s_mid < w || sh.s_mid == 1)) {
        // one distance only: w * h consecutive rows per tile, walked along it
        sh.s_mid = sh.s_big;
        sh.w = w * h;
        sh.h = 1;
        if (sh.w > sh.s_big) sh.s_big = 1;
    }
    sh.slots = 0;                   // the largest tile's distinct columns, rounded up
    return sh;
}

dspmm_status_t get_ring_plan(dspmm_matrix_s *A, int w, int h, int zchunk, RingPlan **out,
                             bool build) {
    const auto key = std::make_tuple(w, h, zchunk);
    auto it = A->ring_plans.find(key);
    if (it != A->ring_plans.end()) { *out = &it->second; return DSPMM_STATUS_SUCCESS; }
    if (!build) return DSPMM_STATUS_NOT_PREPARED;
    const RingShape sh = ring_shape(A, w, h, zchunk);
    RingPlanHost hp;
    RingPlan p;
    p.usable = build_ring_plan(A->h_rowPtr, A->h_colIdx, A->h_values, A->nr, A->nc, sh, hp);
    if (p.usable) {
        p.nstreams = (int)hp.stream_off.size() - 1;
        p.ntiles = (int)hp.tiles.size();
        p.slots = hp.slots;
        p.rows_max = hp.rows_max;
        p.nnz_max = hp.nnz_max;
        p.stream_tiles_max = hp.stream_tiles_max;
        p.stream_fills_max = hp.stream_fills_max;
        p.fill_max = hp.fill_max;
        p.fills = hp.total_fills;
        p.entries = (int64_t)hp.pslot.size();
        // the device copy keeps row ids only (a tile's rows all have its
        // padded length), and one byte per slot where every slot fits
        std::vector<int> row_ids(hp.row_info.size());
        for (size_t i = 0; i < row_ids.size(); i++) row_ids[i] = hp.row_info[i].x;
        p.slot_bytes = hp.zero_slot < 256 ? 1 : 2;
        std::vector<uint8_t> slot8;
        if (p.slot_bytes == 1) slot8.assign(hp.pslot.begin(), hp.pslot.end());
        p.fill_entries = (int64_t)hp.fills.size();
        dspmm_status_t s = upload(&p.d_tiles, hp.tiles.data(), hp.tiles.size(), hipMemcpyHostToDevice, nullptr);
        if (s == DSPMM_STATUS_SUCCESS)
            s = upload(&p.d_streams, hp.stream_off.data(), hp.stream_off.size(), hipMemcpyHostToDevice, nullptr);
        if (s == DSPMM_STATUS_SUCCESS)
            s = upload(&p.d_rows, row_ids.data(), row_ids.size(), hipMemcpyHostToDevice, nullptr);
        if (s == DSPMM_STATUS_SUCCESS)
            s = upload(&p.d_fill_off, hp.fill_off.data(), hp.fill_off.size(), hipMemcpyHostToDevice, nullptr);
        if (s == DSPMM_STATUS_SUCCESS)
            s = upload(&p.d_fills, hp.fills.data(), hp.fills.size(), hipMemcpyHostToDevice, nullptr);
        if (s == DSPMM_STATUS_SUCCESS)
            s = p.slot_bytes == 1
                    ? upload(reinterpret_cast<uint8_t **>(&p.d_pslot), slot8.data(), slot8.size(),
                             hipMemcpyHostToDevice, nullptr)
                    : upload(reinterpret_cast<uint16_t **>(&p.d_pslot), hp.pslot.data(),
                             hp.pslot.size(), hipMemcpyHostToDevice, nullptr);
        if (s == DSPMM_STATUS_SUCCESS)
            s = upload(&p.d_pval, hp.pval.data(), hp.pval.size(), hipMemcpyHostToDevice, nullptr);
        if (s != DSPMM_STATUS_SUCCESS) {
            free_ring_plan(p);
            return s;
        }
        p.bytes = sizeof(int4) * hp.tiles.size() + sizeof(int) * hp.stream_off.size() +
                  sizeof(int) * row_ids.size() + sizeof(int) * hp.fill_off.size() +
                  sizeof(int2) * hp.fills.size() + (size_t)p.slot_bytes * hp.pslot.size() +
                  sizeof(float) * hp.pval.size();
    }
    auto res = A->ring_plans.emplace(key, p);
    *out = &res.first->second;
    return DSPMM_STATUS_SUCCESS;
}

void ring_bind_plan(dspmm_matrix_s *A, const RingPlan &plan, RingLaunch &a) {
    a.tiles = plan.d_tiles;
    a.stream_off = plan.d_streams;
    a.row_info = plan.d_rows;
    a.fill_off = plan.d_fill_off;
    a.fills = plan.d_fills;
    a.pslot = plan.d_pslot;
    a.slot_bytes = plan.slot_bytes;
    a.pval = plan.d_pval;
    a.nstreams = plan.nstreams;
    a.slots = plan.slots;
    a.rows_max = plan.rows_max;
    a.nnz_max = plan.nnz_max;
    a.stream_tiles_max = plan.stream_tiles_max;
    a.stream_fills_max = plan.stream_fills_max;
    a.fill_max = plan.fill_max;
    a.nr = A->nr;
    a.nc = A->nc;
    a.plan_entries = plan.entries;
    a.plan_fills = plan.fill_entries;
}

int ring_auto_z(dspmm_matrix_s *A, int w, int h) {
    const auto key = std::make_pair(w, h);
    auto it = A->ring_auto.find(key);
    if (it != A->ring_auto.end()) return it->second;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int z = ring_auto_chunk(A->nr, ring_shape(A, w, h, 1), cus);
    A->ring_auto.emplace(key, z);
    return z;
}

dspmm_status_t get_plan(dspmm_matrix_s *A, int S, NbPlan **out, int kind, int stride,
                        bool build) {
    const int64_t key = 4 * ((int64_t)S + 16 * (int64_t)(stride - 1)) + kind;
    auto it = A->plans.find(key);
    if (it != A->plans.end()) { *out = &it->second; return DSPMM_STATUS_SUCCESS; }
    if (!build) return DSPMM_STATUS_NOT_PREPARED;
    std::vector<int4> tiles, splits;
    std::vector<int> ucols;
    std::vector<uint16_t> lidx;
    int nslots = 0;
    double reuse = 0;
    if (kind == kPlanSpmv) {
        build_spmv_tiles(A->h_rowPtr, A->nr, S, kSpmvRowMax, tiles, splits, nslots);
    } else if (kind == kPlanCached) {
        build_cached_tiles(A->h_rowPtr, A->h_colIdx, A->nr, A->nc, S, cached_smax(S), tiles, ucols,
                           lidx, reuse);
    } else if (kind == kPlanRowBlock) {
        std::vector<int> boff;
        std::vector<float> rvals;
        std::vector<int2> list;
        dspmm_status_t s = DSPMM_STATUS_SUCCESS;
        NbPlan p;
        p.S = S;
        p.usable = build_row_blocks(A->h_rowPtr, A->h_colIdx, A->h_values, A->nr, S, stride, boff,
                                    list, rvals);
        if (p.usable) {
            p.ntiles = (int)boff.size() - 1;
            p.reuse = list.empty() ? 1.0 : (double)A->nnz / (double)list.size();
            s = upload(&p.d_boff, boff.data(), boff.size(), hipMemcpyHostToDevice, nullptr);
            if (s == DSPMM_STATUS_SUCCESS)
                s = upload(&p.d_list, list.data(), list.size(), hipMemcpyHostToDevice, nullptr);
            if (s == DSPMM_STATUS_SUCCESS)
                s = upload(&p.d_vals, rvals.data(), rvals.size(), hipMemcpyHostToDevice, nullptr);
            if (s != DSPMM_STATUS_SUCCESS) {
                hipFree(p.d_boff);
                hipFree(p.d_list);
                hipFree(p.d_vals);
                return s;
            }
            p.bytes = sizeof(int) * boff.size() + sizeof(int2) * list.size() +
                      sizeof(float) * rvals.size();
        }
        auto res = A->plans.emplace(key, std::move(p));
        *out = &res.first->second;
        return DSPMM_STATUS_SUCCESS;
    } else {
        build_tiles(A->h_rowPtr, A->nr, S, tiles, splits, nslots);
    }
    NbPlan p;
    p.S = S;
    p.ntiles = (int)tiles.size() - (kind == kPlanCached ? 1 : 0);
    p.nsplits = (int)splits.size();
    p.nslots = nslots;
    p.reuse = reuse;
    dspmm_status_t st = upload(&p.d_tiles, tiles.data(), tiles.size(), hipMemcpyHostToDevice, nullptr);
    if (st == DSPMM_STATUS_SUCCESS)
        st = upload(&p.d_splits, splits.data(), splits.size(), hipMemcpyHostToDevice, nullptr);
    if (st == DSPMM_STATUS_SUCCESS)
        st = upload(&p.d_ucols, ucols.data(), ucols.size(), hipMemcpyHostToDevice, nullptr);
    if (st == DSPMM_STATUS_SUCCESS)
        st = upload(&p.d_lidx, lidx.data(), lidx.size(), hipMemcpyHostToDevice, nullptr);
    if (st != DSPMM_STATUS_SUCCESS) {
        hipFree(p.d_tiles);
        hipFree(p.d_splits);
        hipFree(p.d_ucols);
        hipFree(p.d_lidx);
        return st;
    }
    p.bytes = sizeof(int4) * (tiles.size() + splits.size()) + sizeof(int) * ucols.size() +
              sizeof(uint16_t) * lidx.size();
    p.h_tiles.swap(tiles);
    p.h_splits.swap(splits);
    auto res = A->plans.emplace(key, std::move(p));
    *out = &res.first->second;
    return DSPMM_STATUS_SUCCESS;
}

int64_t carry_ld(int nv) { return ((int64_t)nv + 7) / 8 * 8; }


int default_tile(int lanes) { return std::max(16, std::min(256, 16 * lanes)); }

// Fills unset fields of an NNZ_TILED config for these operands.
void complete_nb(dspmm_config &c, int nv, int64_t ldb, int64_t ldc, const void *B, const void *C,
                 bool bf16) {
    const int eb = bf16 ? 2 : 4;
    const int maxvw = bf16 ? 8 : 4;
    const int auto_vw = choose_vw(nv, ldb, ldc, B, C, eb, maxvw);
    if (c.vec <= 0 || c.vec > auto_vw || (auto_vw % c.vec) != 0) c.vec = auto_vw;
    const int need = pow2_ceil(div_up(nv, c.vec));
    if (c.lanes <= 0 || c.lanes > 64 || (c.lanes & (c.lanes - 1)) != 0) c.lanes = std::min(64, need);
    if (c.tile_nnz <= 0) c.tile_nnz = default_tile(c.lanes);
    if (c.unroll <= 0) c.unroll = 8;
}

constexpr int kSpmvTile = 1024;
constexpr int kCachedUmaxDefault = 128;
constexpr double kCachedMinReuse = 1.5;   // sampled nonzeros per distinct column
constexpr double kRowBlockMinReuse = 1.2;
// Column-major work (nnz x nv) from which the static default transposes the
// operands for the row-major nnz-tiled kernel instead of running the
// reference-class column-major kernel: below it, the two transposes cost
// more than they save.
constexpr int64_t kTransposedMinWork = 1 << 20;

dspmm_config heuristic(const dspmm_matrix_s *A, DenseLayout layout, int nv, dspmm_dtype_t bt) {
    dspmm_config c{};
    if (nv == 1) {
        c.variant = DSPMM_VARIANT_SPMV_TILED;
        c.tile_nnz = kSpmvTile;
    } else if (layout == DENSE_ROW_MAJOR || bt == DSPMM_R_16BF ||
               (int64_t)A->nnz * nv >= kTransposedMinWork) {
        c.variant = DSPMM_VARIANT_NNZ_TILED;        // column-major: through transposed copies
    } else {
        c.variant = DSPMM_VARIANT_CSR_SCALAR;
    }
    return c;
}

// Row-major-only native kernels: on column-major operands they run on
// transposed copies of B and C.
bool rm_native(const dspmm_config &c, bool bf16) {
    switch (c.variant) {
    case DSPMM_VARIANT_NNZ_TILED:
    case DSPMM_VARIANT_ROW_BLOCKED:
    case DSPMM_VARIANT_NNZ_CACHED:
    case DSPMM_VARIANT_RING: return true;
    case DSPMM_VARIANT_CSR_SCALAR: return c.unroll > 0 || bf16;
    default: return false;
    }
}

int64_t transposed_ld(int nv) { return ((int64_t)nv + 7) / 8 * 8; }

// Scratch of the transposed path: B^T (nc x ld) then C (nr x ld), each
// rounded up to 256 bytes.
size_t transposed_bytes(const dspmm_matrix_s *A, int nv, bool bf16) {
    const size_t ld = (size_t)transposed_ld(nv);
    const size_t b = (size_t)A->nc * ld * (bf16 ? 2 : 4), c = (size_t)A->nr * ld * 4;
    return (b + 255) / 256 * 256 + (c + 255) / 256 * 256;
}

__global__ void k_zero_2d(float *__restrict__ p, int rows, int cols, int64_t ld) {
    const int64_t n = (int64_t)rows * cols;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t r = i / cols, c = i % cols;
        p[r * ld + c] = 0.f;
    }
}

dspmm_status_t zero_dense(float *C, int rows, int cols, int64_t ld, hipStream_t st) {
    if (rows == 0 || cols == 0) return DSPMM_STATUS_SUCCESS;
    if (ld == cols) return launch_fill_zero(C, sizeof(float) * (size_t)rows * (size_t)cols, st);
    const int64_t n = (int64_t)rows * cols;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_zero_2d, dim3(blocks), dim3(256), 0, st, C, rows, cols, ld);
    return launch_status();
}

// What a run_ex call may do besides launching.
enum RunMode {
    kRunLaunch = 0,     // launch; build missing plans and grow the handle's
                        // buffers first, unless the stream is being captured
    kRunPlans = 1,      // build the plans the call needs; launch nothing
    kRunPrepare = 2,    // build its plans and size the handle's buffers; launch nothing
    kRunNoAlloc = 3,    // launch with everything prepared; allocate and copy nothing
};

bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// The handle's own buffer (carry or transposed copies) of at least `need`
// bytes, grown where the mode allows it.  Sets *buf = nullptr (and
// succeeds) in kRunPlans mode when it is too small: nothing is launched.
dspmm_status_t handle_buffer(void **buf, size_t *have, size_t need, RunMode mode, bool grow,
                             void **out) {
    *out = nullptr;
    if (*have >= need) {
        *out = *buf;
        return DSPMM_STATUS_SUCCESS;
    }
    if (mode == kRunPlans) return DSPMM_STATUS_SUCCESS;
    if (!grow) return DSPMM_STATUS_NOT_PREPARED;
    (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    dspmm_status_t s = hip_status(hipMalloc(buf, need));
    if (s != DSPMM_STATUS_SUCCESS) return s;
    *have = need;
    *out = *buf;
    return DSPMM_STATUS_SUCCESS;
}

// `mid` (may be null) is recorded on the stream after the main kernel of
// the call (before the split-row fix-up of the nnz-tiled variant).
dspmm_status_t run_ex(dspmm_matrix_s *A, DenseLayout layout, int nv, const void *B, int64_t ldb,
                      dspmm_dtype_t bt, float *C, int64_t ldc, const dspmm_config &cfg,
                      void *ws, size_t ws_bytes, hipStream_t st, hipEvent_t mid = nullptr,
                      RunMode mode = kRunLaunch) {
    const bool bf16 = bt == DSPMM_R_16BF;
    const bool launch = mode == kRunLaunch || mode == kRunNoAlloc;
    // plans may be built and the handle's buffers grown here (never under
    // stream capture, where a blocking copy or an allocation would break it)
    const bool lazy = mode == kRunLaunch && !capturing(st);
    const bool build = lazy || mode == kRunPlans || mode == kRunPrepare;
    const bool grow = lazy || mode == kRunPrepare;
    if (layout == DENSE_COL_MAJOR && rm_native(cfg, bf16)) {
        if (nv == 1)      // x and y are contiguous: the row-major call with unit strides
            return run_ex(A, DENSE_ROW_MAJOR, 1, B, 1, bt, C, 1, cfg, ws, ws_bytes, st, mid, mode);
        const size_t tb = transposed_bytes(A, nv, bf16);
        char *base = nullptr;
        if (ws) {
            if (ws_bytes < tb) return DSPMM_STATUS_WORKSPACE_TOO_SMALL;
            if (!aligned(ws, 256)) return DSPMM_STATUS_INVALID_VALUE;
            base = reinterpret_cast<char *>(ws);
        } else {
            void *buf = nullptr;
            dspmm_status_t s = handle_buffer(&A->tbuf, &A->tbuf_bytes, tb, mode, grow, &buf);
            if (s != DSPMM_STATUS_SUCCESS) return s;
            base = reinterpret_cast<char *>(buf);
        }
        const int64_t ld = transposed_ld(nv);
        const size_t bbytes = ((size_t)A->nc * (size_t)ld * (bf16 ? 2 : 4) + 255) / 256 * 256;
        void *Bt = base;
        float *Ct = base ? reinterpret_cast<float *>(base + bbytes) : nullptr;
        dspmm_status_t s = DSPMM_STATUS_SUCCESS;
        // column-major B is nv rows of nc (stride ldb); B^T is nc rows of ld
        if (launch) s = launch_transpose(B, ldb, nv, A->nc, Bt, ld, bf16 ? 2 : 4, st);
        if (s == DSPMM_STATUS_SUCCESS)
            s = run_ex(A, DENSE_ROW_MAJOR, nv, Bt ? Bt : reinterpret_cast<const void *>(256), ld,
                       bt, Ct ? Ct : reinterpret_cast<float *>(256), ld, cfg,
                       ws ? base + tb : nullptr, ws ? ws_bytes - tb : 0, st, mid, mode);
        if (s == DSPMM_STATUS_SUCCESS && launch)
            s = launch_transpose(Ct, ld, A->nr, nv, C, ldc, 4, st);
        return s;
    }
    // Carry buffer of `need` bytes for split rows: the caller's workspace, or
    // the handle's own buffer.
    auto carry_buffer = [&](size_t need, int align, float **carry) -> dspmm_status_t {
        *carry = nullptr;
        if (need == 0) return DSPMM_STATUS_SUCCESS;
        if (ws) {
            if (ws_bytes < need) return DSPMM_STATUS_WORKSPACE_TOO_SMALL;
            if (!aligned(ws, align)) return DSPMM_STATUS_INVALID_VALUE;
            *carry = reinterpret_cast<float *>(ws);
            return DSPMM_STATUS_SUCCESS;
        }
        void *buf = nullptr;
        dspmm_status_t s = handle_buffer(reinterpret_cast<void **>(&A->carry), &A->carry_bytes,
                                         need, mode, grow, &buf);
        *carry = reinterpret_cast<float *>(buf);
        return s;
    };
    switch (cfg.variant) {
    case DSPMM_VARIANT_NNZ_TILED: {
        if (layout != DENSE_ROW_MAJOR) return DSPMM_STATUS_NOT_SUPPORTED;
        dspmm_config c = cfg;
        complete_nb(c, nv, ldb, ldc, B, C, bf16);
        if (nnz_tiled_lds_bytes(c.lanes, c.tile_nnz) > kMaxTileLds) return DSPMM_STATUS_NOT_SUPPORTED;
        NbPlan *plan = nullptr;
        dspmm_status_t s = get_plan(A, c.tile_nnz, &plan, kPlanNnzTiled, 1, build);
        if (s != DSPMM_STATUS_SUCCESS) return s;
        const int64_t ldcar = carry_ld(nv);
        float *carry = nullptr;
        s = carry_buffer(sizeof(float) * (size_t)plan->nslots * (size_t)ldcar, 32, &carry);
        if (s != DSPMM_STATUS_SUCCESS || !launch) return s;
        NbLaunch a;
        ItemDesc &d = a.item;
        d.tiles = plan->d_tiles;
        d.rowPtr = A->rowPtr;
        d.colIdx = A->colIdx;
        d.values = A->values;
        d.B = B;
        d.C = C;
        d.carry = carry;
        d.ldb = ldb;
        d.ldc = ldc;
        d.ldcarry = ldcar;
        d.nv = nv;
        d.slab_w = c.vec * c.lanes;
        d.S = c.tile_nnz;
        d.cls = item_class(c.vec, c.lanes);
        d.ntiles = plan->ntiles;
        d.nslabs = div_up(nv, d.slab_w);
        d.splits = plan->d_splits;
        d.nsplits = plan->nsplits;
        a.ntiles = plan->ntiles;
        a.splits = plan->d_splits;
        a.nsplits = plan->nsplits;
        a.bf16 = bf16;
        a.vw = c.vec;
        a.lanes = c.lanes;
        a.nslabs = d.nslabs;
        a.unroll = c.unroll;
        a.mid_event = mid;
        return launch_nnz_tiled(a, st);
    }
    case DSPMM_VARIANT_NNZ_CACHED: {
        if (layout != DENSE_ROW_MAJOR) return DSPMM_STATUS_NOT_SUPPORTED;
        dspmm_config c = cfg;
        const int umax = c.tile_nnz > 0 ? c.tile_nnz : kCachedUmaxDefault;
        complete_nb(c, nv, ldb, ldc, B, C, bf16);
        if (c.lanes * c.vec < nv) return DSPMM_STATUS_NOT_SUPPORTED;     // one slab only
        if (umax < 8 || umax > 4096 || umax % 8) return DSPMM_STATUS_NOT_SUPPORTED;
        const int smax = cached_smax(umax);
        if (nnz_cached_lds_bytes(umax, smax, nv, c.vec, bf16) > kMaxCachedLds)
            return DSPMM_STATUS_NOT_SUPPORTED;
        NbPlan *plan = nullptr;
        dspmm_status_t s = get_plan(A, umax, &plan, kPlanCached, 1, build);
        if (s != DSPMM_STATUS_SUCCESS || !launch) return s;
        CachedLaunch a;
        a.tiles = plan->d_tiles;
        a.ntiles = plan->ntiles;
        a.ucols = plan->d_ucols;
        a.lidx = plan->d_lidx;
        a.rowPtr = A->rowPtr;
        a.colIdx = A->colIdx;
        a.values = A->values;
        a.B = B;
        a.C = C;
        a.ldb = ldb;
        a.ldc = ldc;
        a.nv = nv;
        a.umax = umax;
        a.smax = smax;
        a.vw = c.vec;
        a.lanes = c.lanes;
        a.bf16 = bf16;
        // workgroup size: cfg.unroll when given (64/128/256), else enough
        // lane groups for the plan's mean rows per tile
        if (cfg.unroll > 0) {
            a.threads = cfg.unroll;
        } else {
            const double rows = plan->ntiles ? (double)A->nr / plan->ntiles : 1.0;
            a.threads = 64;
            while (a.threads < 256 && a.threads < rows * c.lanes) a.threads *= 2;
        }
        if (a.threads != 64 && a.threads != 128 && a.threads != 256) return DSPMM_STATUS_NOT_SUPPORTED;
        if (a.lanes > a.threads) a.threads = a.lanes;
        s = launch_nnz_cached(a, st);
        if (s == DSPMM_STATUS_SUCCESS && mid) s = hip_status(hipEventRecord(mid, st));
        return s;
    }
    case DSPMM_VARIANT_ROW_BLOCKED: {
        if (layout != DENSE_ROW_MAJOR) return DSPMM_STATUS_NOT_SUPPORTED;
        const int R = cfg.tile_nnz > 0 ? cfg.tile_nnz : 4;
        if (R != 2 && R != 4 && R != 8) return DSPMM_STATUS_NOT_SUPPORTED;
        dspmm_config c = cfg;
        c.tile_nnz = 0;
        complete_nb(c, nv, ldb, ldc, B, C, bf16);
        if (cfg.row_stride < -1 || cfg.row_stride > std::max(1, A->nr))
            return DSPMM_STATUS_INVALID_VALUE;
        if (cfg.row_stride < 0 && A->rb_stride < 1 && !build) return DSPMM_STATUS_NOT_PREPARED;
        const int stride = cfg.row_stride < 0 ? shared_row_stride(A) : std::max(1, cfg.row_stride);
        NbPlan *plan = nullptr;
        dspmm_status_t s = get_plan(A, R, &plan, kPlanRowBlock, stride, build);
        if (s != DSPMM_STATUS_SUCCESS) return s;
        if (!plan->usable) return DSPMM_STATUS_NOT_SUPPORTED;
        if (!launch) return DSPMM_STATUS_SUCCESS;
        const Slab sl = make_slab(nv, c.vec);
        RowBlockLaunch a;
        a.boff = plan->d_boff;
        a.list = plan->d_list;
        a.vals = plan->d_vals;
        a.rowPtr = A->rowPtr;
        a.B = B;
        a.C = C;
        a.ldb = ldb;
        a.ldc = ldc;
        a.nr = A->nr;
        a.nv = nv;
        a.nblocks = plan->ntiles;
        a.rows = R;
        a.stride = stride;
        a.vw = c.vec;
        a.lanes = std::min(c.lanes, sl.lanes);
        a.slab_w = a.lanes * c.vec;
        a.nslabs = (nv + a.slab_w - 1) / a.slab_w;
        a.unroll = cfg.unroll > 0 ? cfg.unroll : 8;
        a.bf16 = bf16;
        s = launch_row_blocked(a, st);
        if (s == DSPMM_STATUS_SUCCESS && mid) s = hip_status(hipEventRecord(mid, st));
        return s;
    }
    case DSPMM_VARIANT_RING: {
        // (an empty matrix has nothing to stage: the other kernels take it)
        if (layout != DENSE_ROW_MAJOR || A->nnz == 0) return DSPMM_STATUS_NOT_SUPPORTED;
        const int w = cfg.tile_nnz > 0 ? cfg.tile_nnz : 8;
        const int h = cfg.unroll > 0 ? cfg.unroll : 4;
        if (w > 256 || h > 256 || w * h > 256) return DSPMM_STATUS_NOT_SUPPORTED;
        dspmm_config c = cfg;
        c.tile_nnz = 0;
        c.unroll = 0;
        complete_nb(c, nv, ldb, ldc, B, C, bf16);
        const Slab sl = make_slab(nv, c.vec);
        if (sl.count != 1) return DSPMM_STATUS_NOT_SUPPORTED;          // one slab only
        if (A->rb_stride < 1 && !build) return DSPMM_STATUS_NOT_PREPARED;
        RingLaunch a;
        a.B = B;
        a.C = C;
        a.ldb = ldb;
        a.ldc = ldc;
        a.nv = nv;
        a.vw = c.vec;
        a.lanes = sl.lanes;
        a.bf16 = bf16;
        const int z = cfg.row_stride > 0 ? cfg.row_stride : ring_auto_z(A, w, h);
        RingPlan *plan = nullptr;
        dspmm_status_t s = get_ring_plan(A, w, h, z, &plan, build);
        if (s != DSPMM_STATUS_SUCCESS) return s;
        if (!plan->usable) return DSPMM_STATUS_NOT_SUPPORTED;
        ring_bind_plan(A, *plan, a);
        if (!ring_supported(a)) return DSPMM_STATUS_NOT_SUPPORTED;
        if (!launch) return DSPMM_STATUS_SUCCESS;
        s = launch_ring(a, st);
        if (s == DSPMM_STATUS_SUCCESS && mid) s = hip_status(hipEventRecord(mid, st));
        return s;
    }
    case DSPMM_VARIANT_SPMV_TILED: {
        if (nv != 1) return DSPMM_STATUS_NOT_SUPPORTED;
        const int T = cfg.tile_nnz > 0 ? cfg.tile_nnz : kSpmvTile;
        if (!spmv_tile_ok(T)) return DSPMM_STATUS_NOT_SUPPORTED;
        NbPlan *plan = nullptr;
        dspmm_status_t s = get_plan(A, T, &plan, kPlanSpmv, 1, build);
        if (s != DSPMM_STATUS_SUCCESS) return s;
        const int64_t ldcar = carry_ld(1);
        float *carry = nullptr;
        s = carry_buffer(sizeof(float) * (size_t)plan->nslots * (size_t)ldcar, 4, &carry);
        if (s != DSPMM_STATUS_SUCCESS || !launch) return s;
        NbLaunch a;
        ItemDesc &d = a.item;
        d.tiles = plan->d_tiles;
        d.rowPtr = A->rowPtr;
        d.colIdx = A->colIdx;
        d.values = A->values;
        d.B = B;
        d.C = C;
        d.carry = carry;
        d.ldb = layout == DENSE_ROW_MAJOR ? ldb : 1;   // stride of x
        d.ldc = layout == DENSE_ROW_MAJOR ? ldc : 1;   // stride of y
        d.ldcarry = ldcar;
        d.nv = 1;
        d.slab_w = 1;
        d.S = T;
        d.cls = 0;
        d.ntiles = plan->ntiles;
        d.nslabs = 1;
        d.splits = plan->d_splits;
        d.nsplits = plan->nsplits;
        a.ntiles = plan->ntiles;
        a.splits = plan->d_splits;
        a.nsplits = plan->nsplits;
        a.bf16 = bf16;
        a.mid_event = mid;
        return launch_spmv_tiled(a, st);
    }
    case DSPMM_VARIANT_CSR_SCALAR:
    case DSPMM_VARIANT_CSR_VECTOR:
    case DSPMM_VARIANT_COO_SCALAR:
    case DSPMM_VARIANT_COO_VECTOR: {
        if (cfg.variant == DSPMM_VARIANT_CSR_SCALAR && layout == DENSE_ROW_MAJOR &&
            (bf16 || cfg.unroll > 0)) {
            dspmm_config c = cfg;
            if (cfg.tile_nnz != 0 && cfg.tile_nnz != 1) return DSPMM_STATUS_NOT_SUPPORTED;
            c.tile_nnz = 0;
            c.unroll = cfg.unroll > 0 ? cfg.unroll : 4;
            if (!row_seq_unroll_ok(c.unroll)) return DSPMM_STATUS_NOT_SUPPORTED;
            if (!launch) return DSPMM_STATUS_SUCCESS;
            complete_nb(c, nv, ldb, ldc, B, C, bf16);
            const Slab sl = make_slab(nv, c.vec);
            RowSeqLaunch a;
            a.rowPtr = A->rowPtr;
            a.colIdx = A->colIdx;
            a.values = A->values;
            a.B = B;
            a.C = C;
            a.ldb = ldb;
            a.ldc = ldc;
            a.nr = A->nr;
            a.nv = nv;
            a.vw = c.vec;
            a.lanes = std::min(c.lanes, sl.lanes);
            a.slab_w = a.lanes * c.vec;
            a.nslabs = div_up(nv, a.slab_w);
            a.unroll = c.unroll;
            a.sliced = cfg.tile_nnz == 1;
            a.bf16 = bf16;
            dspmm_status_t s = launch_row_seq(a, st);
            if (s == DSPMM_STATUS_SUCCESS && mid) s = hip_status(hipEventRecord(mid, st));
            return s;
        }
        if (bf16) return DSPMM_STATUS_NOT_SUPPORTED;
        const bool rm = layout == DENSE_ROW_MAJOR;
        int which = 0;
        switch (cfg.variant) {
        case DSPMM_VARIANT_CSR_SCALAR: which = rm ? K_SPMM_CSR_SCALAR_ROW : K_SPMM_CSR_SCALAR_COL; break;
        case DSPMM_VARIANT_CSR_VECTOR: which = rm ? K_SPMM_CSR_VECTOR_ROW : K_SPMM_CSR_VECTOR_COL; break;
        case DSPMM_VARIANT_COO_SCALAR: which = rm ? K_SPMM_COO_SCALAR_ROW : K_SPMM_COO_SCALAR_COL; break;
        default: which = rm ? K_SPMM_COO_VECTOR_ROW : K_SPMM_COO_VECTOR_COL; break;
        }
        const bool coo = cfg.variant == DSPMM_VARIANT_COO_SCALAR || cfg.variant == DSPMM_VARIANT_COO_VECTOR;
        if (coo) {
            if (!A->rowIdx && A->nnz > 0) {
                if (!build) return DSPMM_STATUS_NOT_PREPARED;
                dspmm_status_t s = ensure_rowidx(A);
                if (s != DSPMM_STATUS_SUCCESS) return s;
            }
            if (!launch) return DSPMM_STATUS_SUCCESS;
            dspmm_status_t s = rm ? zero_dense(C, A->nr, nv, ldc, st) : zero_dense(C, nv, A->nr, ldc, st);
            if (s != DSPMM_STATUS_SUCCESS) return s;
        }
        if (!launch) return DSPMM_STATUS_SUCCESS;
        dspmm_status_t s = launch_compat(which, A->nr, A->nc, A->nnz, nv, ldb, ldc, A->rowPtr,
                                         A->rowIdx, A->colIdx, A->values,
                                         reinterpret_cast<const float *>(B), C, st, cfg.lanes);
        if (mid) (void)hipEventRecord(mid, st);
        return s;
    }
    }
    return DSPMM_STATUS_INVALID_VALUE;
}

dspmm_status_t check_ex(dspmm_matrix_s *A, DenseLayout layout, int nv, const void *B, int64_t ldb,
                        dspmm_dtype_t bt, const float *C, int64_t ldc) {
    if (!A || nv < 1) return DSPMM_STATUS_INVALID_VALUE;
    if (layout != DENSE_ROW_MAJOR && layout != DENSE_COL_MAJOR) return DSPMM_STATUS_INVALID_VALUE;
    if (bt != DSPMM_R_32F && bt != DSPMM_R_16BF) return DSPMM_STATUS_INVALID_VALUE;
    if (layout == DENSE_ROW_MAJOR) {
        if (ldb < nv || ldc < nv) return DSPMM_STATUS_INVALID_VALUE;
    } else {
        if (ldb < A->nc || ldc < A->nr) return DSPMM_STATUS_INVALID_VALUE;
    }
    if (A->nr > 0 && !C) return DSPMM_STATUS_INVALID_VALUE;
    if (A->nnz > 0 && !B) return DSPMM_STATUS_INVALID_VALUE;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_config resolve(dspmm_matrix_s *A, DenseLayout layout, int nv, dspmm_dtype_t bt,
                     const dspmm_config *cfg, bool learned) {
    if (cfg && cfg->variant != DSPMM_VARIANT_AUTO) return *cfg;
    const auto key = std::make_tuple((int)layout, nv, (int)bt);
    auto it = A->tuned.find(key);
    if (it != A->tuned.end()) return it->second;
    it = A->prepared.find(key);
    if (it != A->prepared.end()) return it->second;
    dspmm_config c;
    if (learned && layout == DENSE_ROW_MAJOR && learned_select(A, nv, &c)) {
        // the selector was measured with FP32 B; with BF16 B only its B-row
        // ring picks carry over (the ring takes either, with the vector width
        // left to the operands)
        if (bt == DSPMM_R_32F) return c;
        if (c.variant == DSPMM_VARIANT_RING) {
            c.vec = 0;
            c.lanes = 0;
            return c;
        }
    }
    return heuristic(A, layout, nv, bt);
}

// Workspace of a resolved configuration (caller holds A->mu).
dspmm_status_t ex_workspace_bytes(dspmm_matrix_s *A, DenseLayout layout, int nv, dspmm_dtype_t btype,
                                  dspmm_config c, size_t *bytes) {
    *bytes = 0;
    const bool bf16 = btype == DSPMM_R_16BF;
    if (layout == DENSE_COL_MAJOR && rm_native(c, bf16)) {
        size_t inner = 0;
        dspmm_status_t st = ex_workspace_bytes(A, DENSE_ROW_MAJOR, nv, btype, c, &inner);
        if (st != DSPMM_STATUS_SUCCESS) return st;
        *bytes = (nv > 1 ? transposed_bytes(A, nv, bf16) : 0) + inner;
        return DSPMM_STATUS_SUCCESS;
    }
    if (c.variant == DSPMM_VARIANT_SPMV_TILED && nv == 1) {
        NbPlan *plan = nullptr;
        dspmm_status_t st = get_plan(A, c.tile_nnz > 0 ? c.tile_nnz : kSpmvTile, &plan, kPlanSpmv);
        if (st != DSPMM_STATUS_SUCCESS) return st;
        *bytes = sizeof(float) * (size_t)plan->nslots * (size_t)carry_ld(1);
        return DSPMM_STATUS_SUCCESS;
    }
    if (c.variant != DSPMM_VARIANT_NNZ_TILED) return DSPMM_STATUS_SUCCESS;
    // Tile size does not depend on pointers; use aligned placeholders.
    complete_nb(c, nv, nv, nv, reinterpret_cast<const void *>(256), reinterpret_cast<const void *>(256),
                btype == DSPMM_R_16BF);
    NbPlan *plan = nullptr;
    dspmm_status_t st = get_plan(A, c.tile_nnz, &plan);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    *bytes = sizeof(float) * (size_t)plan->nslots * (size_t)carry_ld(nv);
    return DSPMM_STATUS_SUCCESS;
}

float median_of(std::vector<float> v) {
    if (v.empty()) return 0.f;
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

}  // namespace dspmm

// Device scratch of dspmm_spmm_ex_noalloc (split-row carries, transposed
// copies of column-major operands): one copy per GPU, held until the
// process ends.  Host code never uses this name as an address: each
// device's address comes from hipGetSymbolAddress, once, with that device
// current (scratch_address).
static __device__ __align__(256) unsigned char scratch[DSPMM_NOALLOC_SCRATCH_BYTES];

namespace dspmm {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

dspmm_status_t scratch_address(void **out) {
    *out = nullptr;
    int dev = 0;
    dspmm_status_t st = hip_status(hipGetDevice(&dev));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (dev < 0 || dev >= kMaxDevices) return DSPMM_STATUS_NOT_SUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        st = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (st != DSPMM_STATUS_SUCCESS) return st;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return DSPMM_STATUS_SUCCESS;
}
}  // namespace
}  // namespace dspmm

extern "C" {

dspmm_status_t dspmm_fill_zero(void *ptr, size_t bytes, dspmm_stream_t stream) {
    if (bytes && !ptr) return DSPMM_STATUS_INVALID_VALUE;
    return launch_fill_zero(ptr, bytes, as_hip(stream));
}

dspmm_status_t dspmm_convert_f32_to_bf16(const float *src, void *dst, int64_t n,
                                         dspmm_stream_t stream) {
    if (n < 0 || (n > 0 && (!src || !dst))) return DSPMM_STATUS_INVALID_VALUE;
    return launch_f32_to_bf16(src, reinterpret_cast<uint16_t *>(dst), n, as_hip(stream));
}

dspmm_status_t dspmm_matrix_create(dspmm_matrix_t *out, int nr, int nc, int nnz,
                                   const int *h_rowPtr, const int *h_rowIdx, const int *h_colIdx,
                                   const float *h_values, unsigned flags) {
    if (!out) return DSPMM_STATUS_INVALID_VALUE;
    *out = nullptr;
    if (nr < 0 || nc < 0 || nnz < 0) return DSPMM_STATUS_INVALID_VALUE;
    if (nnz > 0 && (!h_colIdx || !h_values)) return DSPMM_STATUS_INVALID_VALUE;
    if (!h_rowPtr && !h_rowIdx && nnz > 0) return DSPMM_STATUS_INVALID_VALUE;
    auto *A = new dspmm_matrix_s;
    A->nr = nr;
    A->nc = nc;
    A->nnz = nnz;
    if (h_rowPtr) {
        A->h_rowPtr.assign(h_rowPtr, h_rowPtr + (size_t)nr + 1);
        if (!rowptr_ok(A->h_rowPtr, nr, nnz)) { delete A; return DSPMM_STATUS_INVALID_VALUE; }
    } else if (!rowptr_from_rows(h_rowIdx, nr, nnz, A->h_rowPtr)) {
        delete A;
        return DSPMM_STATUS_INVALID_VALUE;
    }
    for (int i = 0; i < nnz; i++)
        if (h_colIdx[i] < 0 || h_colIdx[i] >= nc) { delete A; return DSPMM_STATUS_INVALID_VALUE; }
    A->h_colIdx.assign(h_colIdx, h_colIdx + nnz);
    A->h_values.assign(h_values, h_values + nnz);
    compute_stats(A);
    dspmm_status_t st = upload(&A->rowPtr, A->h_rowPtr.data(), (size_t)nr + 1, hipMemcpyHostToDevice, nullptr);
    if (st == DSPMM_STATUS_SUCCESS)
        st = upload(&A->colIdx, h_colIdx, (size_t)nnz, hipMemcpyHostToDevice, nullptr);
    if (st == DSPMM_STATUS_SUCCESS)
        st = upload(&A->values, h_values, (size_t)nnz, hipMemcpyHostToDevice, nullptr);
    if (st == DSPMM_STATUS_SUCCESS && (flags & DSPMM_MATRIX_WITH_COO)) {
        if (h_rowIdx) st = upload(&A->rowIdx, h_rowIdx, (size_t)nnz, hipMemcpyHostToDevice, nullptr);
        else st = ensure_rowidx(A);
    }
    if (st != DSPMM_STATUS_SUCCESS) { free_matrix(A); return st; }
    A->base_bytes = sizeof(int) * ((size_t)nr + 1 + (size_t)nnz) + sizeof(float) * (size_t)nnz;
    *out = A;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_matrix_create_device(dspmm_matrix_t *out, int nr, int nc, int nnz,
                                          const int *d_rowPtr, const int *d_rowIdx,
                                          const int *d_colIdx, const float *d_values,
                                          unsigned flags, dspmm_stream_t stream) {
    if (!out) return DSPMM_STATUS_INVALID_VALUE;
    *out = nullptr;
    if (nr < 0 || nc < 0 || nnz < 0 || !d_rowPtr) return DSPMM_STATUS_INVALID_VALUE;
    if (nnz > 0 && (!d_colIdx || !d_values)) return DSPMM_STATUS_INVALID_VALUE;
    std::vector<int> rp((size_t)nr + 1), ci((size_t)nnz);
    std::vector<float> va((size_t)nnz);
    dspmm_status_t st = hip_status(hipStreamSynchronize(as_hip(stream)));
    // host copies, once: plans are built from them
    if (st == DSPMM_STATUS_SUCCESS)
        st = hip_status(hipMemcpy(rp.data(), d_rowPtr, sizeof(int) * rp.size(), hipMemcpyDeviceToHost));
    if (st == DSPMM_STATUS_SUCCESS && nnz)
        st = hip_status(hipMemcpy(ci.data(), d_colIdx, sizeof(int) * ci.size(), hipMemcpyDeviceToHost));
    if (st == DSPMM_STATUS_SUCCESS && nnz)
        st = hip_status(hipMemcpy(va.data(), d_values, sizeof(float) * va.size(), hipMemcpyDeviceToHost));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (!rowptr_ok(rp, nr, nnz)) return DSPMM_STATUS_INVALID_VALUE;
    for (int i = 0; i < nnz; i++)
        if (ci[(size_t)i] < 0 || ci[(size_t)i] >= nc) return DSPMM_STATUS_INVALID_VALUE;
    auto *A = new dspmm_matrix_s;
    A->nr = nr;
    A->nc = nc;
    A->nnz = nnz;
    A->h_rowPtr.swap(rp);
    A->h_colIdx.swap(ci);
    A->h_values.swap(va);
    compute_stats(A);
    const hipStream_t s = as_hip(stream);
    st = upload(&A->rowPtr, d_rowPtr, (size_t)nr + 1, hipMemcpyDeviceToDevice, s);
    if (st == DSPMM_STATUS_SUCCESS) st = upload(&A->colIdx, d_colIdx, (size_t)nnz, hipMemcpyDeviceToDevice, s);
    if (st == DSPMM_STATUS_SUCCESS) st = upload(&A->values, d_values, (size_t)nnz, hipMemcpyDeviceToDevice, s);
    if (st == DSPMM_STATUS_SUCCESS && (flags & DSPMM_MATRIX_WITH_COO)) {
        if (d_rowIdx) st = upload(&A->rowIdx, d_rowIdx, (size_t)nnz, hipMemcpyDeviceToDevice, s);
        else st = ensure_rowidx(A);
    }
    if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipStreamSynchronize(s));
    if (st != DSPMM_STATUS_SUCCESS) { free_matrix(A); return st; }
    A->base_bytes = sizeof(int) * ((size_t)nr + 1 + (size_t)nnz) + sizeof(float) * (size_t)nnz;
    *out = A;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_matrix_destroy(dspmm_matrix_t A) {
    if (!A) return DSPMM_STATUS_SUCCESS;
    free_matrix(A);       // the caller guarantees no launch still reads the handle
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_matrix_prepare(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                                    int64_t ldb, dspmm_dtype_t btype, const float *C, int64_t ldc,
                                    const dspmm_config *cfg, unsigned flags) {
    dspmm_status_t st = check_ex(A, layout, nv, B, ldb, btype, C, ldc);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (flags & ~(unsigned)DSPMM_PREPARE_HANDLE_BUFFERS) return DSPMM_STATUS_INVALID_VALUE;
    if (A->nr == 0) return DSPMM_STATUS_SUCCESS;
    std::lock_guard<std::mutex> lk(A->mu);
    // sampled features (host only) computed now, not during a launch
    (void)column_reuse(A);
    (void)shared_row_stride(A);
    const bool auto_cfg = !cfg || cfg->variant == DSPMM_VARIANT_AUTO;
    const auto key = std::make_tuple((int)layout, nv, (int)btype);
    const dspmm_config c = resolve(A, layout, nv, btype, cfg);
    if (auto_cfg && !A->tuned.count(key)) A->prepared[key] = c;
    const RunMode mode = (flags & DSPMM_PREPARE_HANDLE_BUFFERS) ? kRunPrepare : kRunPlans;
    float *Cw = const_cast<float *>(C);
    st = run_ex(A, layout, nv, B, ldb, btype, Cw, ldc, c, nullptr, 0, nullptr, nullptr, mode);
    if (auto_cfg) {
        // the static default AUTO falls back to when the pick cannot run
        const dspmm_status_t sh = run_ex(A, layout, nv, B, ldb, btype, Cw, ldc,
                                         heuristic(A, layout, nv, btype), nullptr, 0, nullptr,
                                         nullptr, mode);
        if (st == DSPMM_STATUS_NOT_SUPPORTED) st = sh;
    }
    return st;
}

dspmm_status_t dspmm_matrix_describe(dspmm_matrix_t A, dspmm_matrix_desc *d) {
    if (!A || !d) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(A->mu);
    d->nr = A->nr;
    d->nc = A->nc;
    d->nnz = A->nnz;
    d->rowPtr = A->rowPtr;
    d->rowIdx = A->rowIdx;
    d->colIdx = A->colIdx;
    d->values = A->values;
    d->hbm_bytes = A->hbm_bytes();
    d->max_row_len = A->max_len;
    d->mean_row_len = A->mean;
    d->cv_row_len = A->cv;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_spmm_ex_resolve(dspmm_matrix_t A, DenseLayout layout, int nv,
                                     dspmm_dtype_t btype, const dspmm_config *cfg,
                                     dspmm_config *resolved) {
    if (!A || !resolved || nv < 1) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(A->mu);
    *resolved = resolve(A, layout, nv, btype, cfg);
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_spmm_ex_workspace_size(dspmm_matrix_t A, DenseLayout layout, int nv,
                                            dspmm_dtype_t btype, const dspmm_config *cfg,
                                            size_t *bytes) {
    if (!A || !bytes || nv < 1) return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(A->mu);
    dspmm_status_t st = ex_workspace_bytes(A, layout, nv, btype, resolve(A, layout, nv, btype, cfg),
                                           bytes);
    if (st != DSPMM_STATUS_SUCCESS || (cfg && cfg->variant != DSPMM_VARIANT_AUTO)) return st;
    // AUTO may fall back to the static default when the resolved
    // configuration cannot run these operands: cover its workspace too.
    size_t h = 0;
    st = ex_workspace_bytes(A, layout, nv, btype, heuristic(A, layout, nv, btype), &h);
    *bytes = std::max(*bytes, h);
    return st;
}
dspmm_status_t dspmm_spmm_ex(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                             int64_t ldb, dspmm_dtype_t btype, float *C, int64_t ldc,
                             const dspmm_config *cfg, void *workspace, size_t workspace_bytes,
                             dspmm_stream_t stream) {
    dspmm_status_t st = check_ex(A, layout, nv, B, ldb, btype, C, ldc);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (A->nr == 0) return DSPMM_STATUS_SUCCESS;
    std::lock_guard<std::mutex> lk(A->mu);
    const dspmm_config c = resolve(A, layout, nv, btype, cfg);
    st = run_ex(A, layout, nv, B, ldb, btype, C, ldc, c, workspace, workspace_bytes, as_hip(stream));
    if (st == DSPMM_STATUS_NOT_SUPPORTED && (!cfg || cfg->variant == DSPMM_VARIANT_AUTO)) {
        // a selected configuration these operands cannot run (alignment,
        // LDS): the static default instead
        const dspmm_config h = heuristic(A, layout, nv, btype);
        st = run_ex(A, layout, nv, B, ldb, btype, C, ldc, h, workspace, workspace_bytes,
                    as_hip(stream));
    }
    return st;
}

dspmm_status_t dspmm_spmm_ex_noalloc(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                                     int64_t ldb, dspmm_dtype_t btype, float *C, int64_t ldc,
                                     const dspmm_config *cfg, dspmm_stream_t stream) {
    dspmm_status_t st = check_ex(A, layout, nv, B, ldb, btype, C, ldc);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (A->nr == 0) return DSPMM_STATUS_SUCCESS;
    void *ws = nullptr;
    st = scratch_address(&ws);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    std::lock_guard<std::mutex> lk(A->mu);
    const bool auto_cfg = !cfg || cfg->variant == DSPMM_VARIANT_AUTO;
    dspmm_config c{};
    if (auto_cfg) {
        // only a configuration already resolved (tuned, or picked by
        // dspmm_matrix_prepare): resolving afresh would compute features
        const auto key = std::make_tuple((int)layout, nv, (int)btype);
        auto it = A->tuned.find(key);
        if (it != A->tuned.end()) {
            c = it->second;
        } else {
            auto ip = A->prepared.find(key);
            if (ip == A->prepared.end()) return DSPMM_STATUS_NOT_PREPARED;
            c = ip->second;
        }
    } else {
        c = *cfg;
    }
    const hipStream_t hs = as_hip(stream);
    st = run_ex(A, layout, nv, B, ldb, btype, C, ldc, c, ws, DSPMM_NOALLOC_SCRATCH_BYTES, hs,
                nullptr, kRunNoAlloc);
    if (st == DSPMM_STATUS_NOT_SUPPORTED && auto_cfg)
        st = run_ex(A, layout, nv, B, ldb, btype, C, ldc, heuristic(A, layout, nv, btype), ws,
                    DSPMM_NOALLOC_SCRATCH_BYTES, hs, nullptr, kRunNoAlloc);
    return st;
}

dspmm_status_t dspmm_benchmark_ex(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                                  int64_t ldb, dspmm_dtype_t btype, float *C, int64_t ldc,
                                  const dspmm_config *cfg, int warmup, int reps, float *ms_mean,
                                  float *ms_per_launch, float *ms_main_per_launch,
                                  dspmm_stream_t stream) {
    if (reps < 1 || warmup < 0) return DSPMM_STATUS_INVALID_VALUE;
    dspmm_status_t st = check_ex(A, layout, nv, B, ldb, btype, C, ldc);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (A->nr == 0) return DSPMM_STATUS_INVALID_VALUE;
    const hipStream_t hs = as_hip(stream);
    dspmm_config c;
    {
        std::lock_guard<std::mutex> lk(A->mu);
        c = resolve(A, layout, nv, btype, cfg);
    }
    for (int i = 0; i < warmup && st == DSPMM_STATUS_SUCCESS; i++) {
        std::lock_guard<std::mutex> lk(A->mu);
        st = run_ex(A, layout, nv, B, ldb, btype, C, ldc, c, nullptr, 0, hs);
    }
    if (st != DSPMM_STATUS_SUCCESS) return st;
    st = hip_status(hipStreamSynchronize(hs));
    if (st != DSPMM_STATUS_SUCCESS) return st;
    // Events only where an output asks for them: between launches for
    // ms_per_launch, after the main kernel for ms_main_per_launch, and
    // otherwise just around the whole timed region.
    const bool per = ms_per_launch || ms_main_per_launch;
    std::vector<hipEvent_t> ev(per ? (size_t)reps + 1 : 2, nullptr);
    std::vector<hipEvent_t> mid(ms_main_per_launch ? (size_t)reps : 0, nullptr);
    for (auto &e : ev) if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventCreate(&e));
    for (auto &e : mid) if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventCreate(&e));
    if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventRecord(ev[0], hs));
    for (int i = 0; i < reps && st == DSPMM_STATUS_SUCCESS; i++) {
        {
            std::lock_guard<std::mutex> lk(A->mu);
            st = run_ex(A, layout, nv, B, ldb, btype, C, ldc, c, nullptr, 0, hs,
                        mid.empty() ? nullptr : mid[(size_t)i]);
        }
        if (st == DSPMM_STATUS_SUCCESS && (per || i == reps - 1))
            st = hip_status(hipEventRecord(ev[per ? (size_t)i + 1 : 1], hs));
    }
    if (st == DSPMM_STATUS_SUCCESS) st = hip_status(hipEventSynchronize(ev.back()));
    if (st == DSPMM_STATUS_SUCCESS) {
        float total = 0;
        st = hip_status(hipEventElapsedTime(&total, ev.front(), ev.back()));
        for (int i = 0; per && i < reps && st == DSPMM_STATUS_SUCCESS; i++) {
            float t = 0, tm = 0;
            st = hip_status(hipEventElapsedTime(&t, ev[(size_t)i], ev[(size_t)i + 1]));
            if (st == DSPMM_STATUS_SUCCESS && !mid.empty())
                st = hip_status(hipEventElapsedTime(&tm, ev[(size_t)i], mid[(size_t)i]));
            if (ms_per_launch) ms_per_launch[i] = t;
            if (ms_main_per_launch) ms_main_per_launch[i] = tm;
        }
        if (ms_mean) *ms_mean = total / (float)reps;
    }
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    for (auto &e : mid) if (e) (void)hipEventDestroy(e);
    return st;
}

}  // extern "C"

namespace {

// Candidate list of the autotuner: the native nnz-tiled kernel over vector
// widths (each with the lane count that covers nv in one slab), tile sizes
// around the default and both unrolls; the row-sequential kernel over its
// unrolls; the row kernels also in column slabs (all of these also for
// column-major operands, on transposed copies); the SpMV tiles when
// nv == 1; then the four reference algorithm classes (FP32 B only).
std::vector<dspmm_config> autotune_candidates(dspmm_matrix_s *A, DenseLayout layout, int nv,
                                              int64_t ldb, int64_t ldc, const void *B,
                                              const float *C, bool bf16) {
    std::vector<dspmm_config> cand;
    // Column-major operands run the row-major kernels on transposed copies
    // (leading dimension transposed_ld, 256-byte aligned), so every row-major
    // candidate applies with those operands.
    const bool cm = layout == DENSE_COL_MAJOR;
    const void *al = reinterpret_cast<const void *>(256);
    if (cm) {
        ldb = ldc = nv > 1 ? transposed_ld(nv) : 1;
        B = al;
        C = reinterpret_cast<const float *>(al);
    }
    // Lane counts of the nnz-tiled and row-sequential kernels: one slab
    // covering nv (0 = auto), and half / quarter slabs, which run the columns
    // in passes (grid.y is the slab, so one pass over all rows ends before
    // the next begins) and shrink the B footprint a pass keeps in L2 at the
    // cost of re-reading A.
    std::vector<int> slab_lanes{0};
    if (nv > 1) {
        dspmm_config base{};
        base.variant = DSPMM_VARIANT_NNZ_TILED;
        complete_nb(base, nv, ldb, ldc, B, C, bf16);
        for (int l = base.lanes / 2; l >= 1 && l * 4 >= base.lanes; l /= 2) slab_lanes.push_back(l);
    }
    if (nv > 1) {
        // column-cached tiles, only where neighbouring rows share columns
        double reuse;
        {
            std::lock_guard<std::mutex> lk(A->mu);
            reuse = column_reuse(A);
        }
        int stride;
        {
            std::lock_guard<std::mutex> lk(A->mu);
            stride = shared_row_stride(A);
        }
        std::vector<int> strides;
        if (reuse >= kRowBlockMinReuse) strides.push_back(1);
        if (stride > 1) strides.push_back(stride);
        for (int rs : strides)
            for (int R : {2, 4, 8})
                for (int u : {4, 8}) {
                    dspmm_config c{};
                    c.variant = DSPMM_VARIANT_ROW_BLOCKED;
                    c.tile_nnz = R;
                    c.unroll = u;
                    c.row_stride = rs;
                    cand.push_back(c);
                }
        if (stride > 1 || reuse >= kRowBlockMinReuse) {
            // B-row ring in LDS: tile shapes (w x h) and tiles per walk:
            // the library's length for this GPU's CUs (0) and half of it
            // (twice the walks, so the two workgroups of a CU start them
            // out of step; measured better where one walk per workgroup
            // slot would start them all at once)
            for (const auto &wh : {std::make_pair(8, 4), std::make_pair(8, 8),
                                   std::make_pair(16, 4), std::make_pair(4, 8)}) {
                int half;
                {
                    std::lock_guard<std::mutex> lk(A->mu);
                    half = ring_auto_z(A, wh.first, wh.second) / 2;
                }
                const int zs[2] = {0, half};
                for (int q = 0; q < (half >= 1 ? 2 : 1); q++) {
                    const int z = zs[q];
                    dspmm_config c{};
                    c.variant = DSPMM_VARIANT_RING;
                    c.tile_nnz = wh.first;
                    c.unroll = wh.second;
                    c.row_stride = z;
                    cand.push_back(c);
                }
            }
        }
        if (reuse >= kCachedMinReuse) {
            dspmm_config base{};
            base.variant = DSPMM_VARIANT_NNZ_TILED;
            complete_nb(base, nv, ldb, ldc, B, C, bf16);
            for (int vw = base.vec; vw >= 1 && vw * 2 >= base.vec; vw /= 2) {
                for (int umax : {64, 128, 256, 512}) {
                    for (int nt : {64, 128, 256}) {
                        dspmm_config c{};
                        c.variant = DSPMM_VARIANT_NNZ_CACHED;
                        c.vec = vw;
                        c.tile_nnz = umax;
                        c.unroll = nt;
                        cand.push_back(c);
                    }
                }
            }
        }
    }
    {
        dspmm_config base{};
        base.variant = DSPMM_VARIANT_NNZ_TILED;
        complete_nb(base, nv, ldb, ldc, B, C, bf16);
        for (int vw = base.vec; vw >= 1; vw /= 2) {
            dspmm_config c{};
            c.variant = DSPMM_VARIANT_NNZ_TILED;
            c.vec = vw;
            complete_nb(c, nv, ldb, ldc, B, C, bf16);
            if (c.vec != vw) continue;
            const int s0 = default_tile(c.lanes);
            for (int S : {s0 / 8, s0 / 4, s0 / 2, s0, s0 * 2, s0 * 4}) {
                if (S < 8 || S > 2048) continue;
                if (nnz_tiled_lds_bytes(c.lanes, S) > kMaxTileLds) continue;
                for (int u : {4, 8}) {
                    dspmm_config x = c;
                    x.tile_nnz = S;
                    x.unroll = u;
                    cand.push_back(x);
                }
            }
            if (vw == base.vec && slab_lanes.size() > 1) {   // half slabs, default tiles
                dspmm_config x = c;
                x.lanes = slab_lanes[1];
                x.tile_nnz = default_tile(x.lanes);
                if (nnz_tiled_lds_bytes(x.lanes, x.tile_nnz) <= kMaxTileLds)
                    for (int u : {4, 8}) {
                        x.unroll = u;
                        cand.push_back(x);
                    }
            }
        }
    }
    {
        // row-sequential (LDS-staged indices), over the gathers in flight;
        // sliced staging windows where some wavefront's rows overflow one
        // window (otherwise both stagings run the same code)
        for (int u : {2, 4, 8, 16})
            for (int lanes : slab_lanes) {
                if (lanes && (u == 2 || u == 16)) continue;
                dspmm_config c{};
                c.variant = DSPMM_VARIANT_CSR_SCALAR;
                c.unroll = u;
                c.lanes = lanes;
                cand.push_back(c);
            }
        dspmm_config base{};
        base.variant = DSPMM_VARIANT_NNZ_TILED;
        complete_nb(base, nv, ldb, ldc, B, C, bf16);
        const int g = 64 / std::min(base.lanes, make_slab(nv, base.vec).lanes);   // rows per wavefront
        bool overflow = false;
        for (int rw = 0; rw < A->nr && !overflow; rw += g)
            overflow = A->h_rowPtr[(size_t)std::min(A->nr, rw + g)] - A->h_rowPtr[(size_t)rw] >
                       kRowSeqWindow;
        if (overflow)
            for (int u : {4, 8, 16}) {
                dspmm_config c{};
                c.variant = DSPMM_VARIANT_CSR_SCALAR;
                c.unroll = u;
                c.tile_nnz = 1;
                cand.push_back(c);
            }
    }
    if (nv == 1) {
        for (int T : {512, 1024, 2048, 4096}) {
            dspmm_config c{};
            c.variant = DSPMM_VARIANT_SPMV_TILED;
            c.tile_nnz = T;
            cand.push_back(c);
        }
    }
    if (!bf16) {
        for (int v : {DSPMM_VARIANT_CSR_SCALAR, DSPMM_VARIANT_CSR_VECTOR, DSPMM_VARIANT_COO_SCALAR,
                      DSPMM_VARIANT_COO_VECTOR}) {
            dspmm_config c{};
            c.variant = v;
            cand.push_back(c);
        }
    }
    return cand;
}

// Median ms of each candidate over `reps` timed calls (after one warm-up);
// a negative value marks a candidate these operands do not support.
// Front runners timed a second time by dspmm_autotune.
constexpr int kTuneFront = 4;

dspmm_status_t time_candidates(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                               int64_t ldb, dspmm_dtype_t btype, float *C, int64_t ldc, int reps,
                               const std::vector<dspmm_config> &cand, std::vector<float> &med,
                               dspmm_stream_t stream) {
    std::vector<float> ms((size_t)reps);
    med.assign(cand.size(), -1.0f);
    for (size_t i = 0; i < cand.size(); i++) {
        float mean = 0;
        const dspmm_status_t st = dspmm_benchmark_ex(A, layout, nv, B, ldb, btype, C, ldc,
                                                     &cand[i], 1, reps, &mean, ms.data(), nullptr,
                                                     stream);
        if (st == DSPMM_STATUS_NOT_SUPPORTED) continue;
        if (st != DSPMM_STATUS_SUCCESS) return st;
        med[i] = median_of(ms);
    }
    return DSPMM_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

dspmm_status_t dspmm_autotune(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                              int64_t ldb, dspmm_dtype_t btype, float *C, int64_t ldc, int reps,
                              dspmm_config *best, float *best_ms, dspmm_stream_t stream) {
    dspmm_status_t st = check_ex(A, layout, nv, B, ldb, btype, C, ldc);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (reps < 1) reps = 5;
    const std::vector<dspmm_config> cand =
        autotune_candidates(A, layout, nv, ldb, ldc, B, C, btype == DSPMM_R_16BF);
    std::vector<float> med;
    st = time_candidates(A, layout, nv, B, ldb, btype, C, ldc, reps, cand, med, stream);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    float best_t = 1e30f;
    dspmm_config best_c;
    {
        std::lock_guard<std::mutex> lk(A->mu);
        best_c = heuristic(A, layout, nv, btype);
    }
    for (size_t i = 0; i < cand.size(); i++)
        if (med[i] >= 0 && med[i] < best_t) {
            best_t = med[i];
            best_c = cand[i];
        }
    // The front runners (within 5% of the best, at most kTuneFront) are
    // timed again with four times the calls, in both orders, and the better
    // of each one's two medians decides: one short timing each picks at
    // random among configurations a few percent apart.
    std::vector<size_t> front;
    for (size_t i = 0; i < cand.size(); i++)
        if (med[i] >= 0 && med[i] <= 1.05f * best_t) front.push_back(i);
    std::sort(front.begin(), front.end(), [&](size_t x, size_t y) { return med[x] < med[y]; });
    if (front.size() > (size_t)kTuneFront) front.resize((size_t)kTuneFront);
    if (front.size() > 1) {
        std::vector<dspmm_config> sub;
        for (size_t i : front) sub.push_back(cand[i]);
        std::vector<float> t1, t2;
        st = time_candidates(A, layout, nv, B, ldb, btype, C, ldc, 4 * reps, sub, t1, stream);
        if (st != DSPMM_STATUS_SUCCESS) return st;
        std::reverse(sub.begin(), sub.end());
        st = time_candidates(A, layout, nv, B, ldb, btype, C, ldc, 4 * reps, sub, t2, stream);
        if (st != DSPMM_STATUS_SUCCESS) return st;
        std::reverse(sub.begin(), sub.end());
        std::reverse(t2.begin(), t2.end());
        float bt = 1e30f;
        for (size_t k = 0; k < sub.size(); k++) {
            float t = t1[k] >= 0 ? t1[k] : 1e30f;
            if (t2[k] >= 0) t = std::min(t, t2[k]);
            if (t < bt) {
                bt = t;
                best_c = sub[k];
            }
        }
        if (bt < 1e30f) best_t = bt;
    }
    {
        std::lock_guard<std::mutex> lk(A->mu);
        A->tuned[std::make_tuple((int)layout, nv, (int)btype)] = best_c;
    }
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_matrix_set_config(dspmm_matrix_t A, DenseLayout layout, int nv,
                                       dspmm_dtype_t btype, const dspmm_config *cfg) {
    if (!A || !cfg || nv < 1 || (layout != DENSE_ROW_MAJOR && layout != DENSE_COL_MAJOR) ||
        (btype != DSPMM_R_32F && btype != DSPMM_R_16BF) || cfg->variant < DSPMM_VARIANT_AUTO ||
        cfg->variant > DSPMM_VARIANT_RING)
        return DSPMM_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(A->mu);
    const auto key = std::make_tuple((int)layout, nv, (int)btype);
    if (cfg->variant == DSPMM_VARIANT_AUTO) A->tuned.erase(key);
    else A->tuned[key] = *cfg;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_autotune_table(dspmm_matrix_t A, DenseLayout layout, int nv, const void *B,
                                    int64_t ldb, dspmm_dtype_t btype, float *C, int64_t ldc,
                                    int reps, dspmm_config *configs, float *ms, int capacity,
                                    int *count, dspmm_stream_t stream) {
    if (!count || capacity < 0 || (capacity > 0 && (!configs || !ms)))
        return DSPMM_STATUS_INVALID_VALUE;
    dspmm_status_t st = check_ex(A, layout, nv, B, ldb, btype, C, ldc);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    if (reps < 1) reps = 5;
    const std::vector<dspmm_config> cand =
        autotune_candidates(A, layout, nv, ldb, ldc, B, C, btype == DSPMM_R_16BF);
    *count = (int)cand.size();
    if (capacity == 0) return DSPMM_STATUS_SUCCESS;
    if (capacity < (int)cand.size()) return DSPMM_STATUS_WORKSPACE_TOO_SMALL;
    std::vector<float> med;
    st = time_candidates(A, layout, nv, B, ldb, btype, C, ldc, reps, cand, med, stream);
    if (st != DSPMM_STATUS_SUCCESS) return st;
    for (size_t i = 0; i < cand.size(); i++) {
        configs[i] = cand[i];
        ms[i] = med[i];
    }
    return DSPMM_STATUS_SUCCESS;
}

}  // extern "C"
