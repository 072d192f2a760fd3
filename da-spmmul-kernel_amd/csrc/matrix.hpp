// matrix.hpp -- the device-resident CSR handle and its nnz-tile plans
// (shared by the single-matrix and batched native paths).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.hpp"

namespace dspmm {

struct NbPlan {
    int S = 0;
    int ntiles = 0, nsplits = 0, nslots = 0;
    int4 *d_tiles = nullptr;
    int4 *d_splits = nullptr;         // {0, row, first_slot, nchunks}
    std::vector<int4> h_tiles, h_splits;
    size_t bytes = 0;
};

// LDS budget of one nnz-tiled workgroup (keeps two workgroups per CU).
constexpr size_t kMaxTileLds = 64 * 1024;

void build_tiles(const std::vector<int> &rp, int nr, int S, std::vector<int4> &tiles,
                 std::vector<int4> &splits, int &nslots);
// SpMV plan: tiles of <= T nonzeros and <= 256 rows, each row <= row_max
// long; a longer row gets its own tile(s) of <= T nonzeros.
void build_spmv_tiles(const std::vector<int> &rp, int nr, int T, int row_max,
                      std::vector<int4> &tiles, std::vector<int4> &splits, int &nslots);

enum PlanKind { kPlanNnzTiled = 0, kPlanSpmv = 1 };

}  // namespace dspmm

struct dspmm_matrix_s {
    int nr = 0, nc = 0, nnz = 0;
    int *rowPtr = nullptr, *rowIdx = nullptr, *colIdx = nullptr;
    float *values = nullptr;
    std::vector<int> h_rowPtr;
    int max_len = 0;
    double mean = 0, cv = 0;
    size_t base_bytes = 0;
    std::mutex mu;
    std::map<int, dspmm::NbPlan> plans;                            // by 2 * tile size + kind
    std::map<std::tuple<int, int, int>, dspmm_config> tuned;       // (layout, nv, btype)
    float *carry = nullptr;
    size_t carry_bytes = 0;

    size_t hbm_bytes() {
        size_t b = base_bytes + carry_bytes;
        if (rowIdx) b += sizeof(int) * (size_t)nnz;
        for (auto &kv : plans) b += kv.second.bytes;
        return b;
    }
};

namespace dspmm {

// All of these expect the caller to hold A->mu.
dspmm_status_t get_plan(dspmm_matrix_s *A, int S, NbPlan **out, int kind = kPlanNnzTiled);
int64_t carry_ld(int nv);
void complete_nb(dspmm_config &c, int nv, int64_t ldb, int64_t ldc, const void *B, const void *C,
                 bool bf16);
dspmm_config resolve(dspmm_matrix_s *A, DenseLayout layout, int nv, dspmm_dtype_t bt,
                     const dspmm_config *cfg);

}  // namespace dspmm
