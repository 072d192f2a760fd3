// rocsparse_cmp.cpp -- rocSPARSE generic SpMV/SpMM as the comparator columns
// of the sweep (the role cuSPARSE has in cusparse_helper.h:120-437).
// Built into lib/libdspmm_rocsparse.so; the kernel library does not link it.
//
// Per (nv, column) the handle keeps a prepared call: dense descriptors, the
// workspace (allocated once; the reference allocates and frees it around
// every call, cusparse_helper.h:220-230) and the preprocess stage done, so
// timed calls run the compute stage only.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <map>
#include <utility>

#include "dspmm_rocsparse.h"

namespace {

struct Column {
    bool csr;
    int layout;
    int alg;  // rocsparse_spmv_alg for nv == 1, rocsparse_spmm_alg otherwise
    const char *name;
};

const Column kSpmv[4] = {
    {true, DENSE_ROW_MAJOR, rocsparse_spmv_alg_default, "csr_default"},
    {true, DENSE_ROW_MAJOR, rocsparse_spmv_alg_csr_adaptive, "csr_adaptive"},
    {false, DENSE_ROW_MAJOR, rocsparse_spmv_alg_default, "coo_default"},
    {false, DENSE_ROW_MAJOR, rocsparse_spmv_alg_coo_atomic, "coo_atomic"},
};
const Column kSpmm[10] = {
    {true, DENSE_ROW_MAJOR, rocsparse_spmm_alg_default, "csr_row_default"},
    {true, DENSE_ROW_MAJOR, rocsparse_spmm_alg_csr_row_split, "csr_row_row_split"},
    {false, DENSE_ROW_MAJOR, rocsparse_spmm_alg_default, "coo_row_default"},
    {false, DENSE_ROW_MAJOR, rocsparse_spmm_alg_coo_segmented_atomic, "coo_row_segmented_atomic"},
    {true, DENSE_COL_MAJOR, rocsparse_spmm_alg_default, "csr_col_default"},
    {true, DENSE_COL_MAJOR, rocsparse_spmm_alg_csr, "csr_col_csr"},
    {false, DENSE_COL_MAJOR, rocsparse_spmm_alg_default, "coo_col_default"},
    {false, DENSE_COL_MAJOR, rocsparse_spmm_alg_coo_atomic, "coo_col_atomic"},
    {false, DENSE_COL_MAJOR, rocsparse_spmm_alg_coo_segmented, "coo_col_segmented"},
    {false, DENSE_COL_MAJOR, rocsparse_spmm_alg_coo_segmented_atomic, "coo_col_segmented_atomic"},
};

const Column *column_of(int nv, int c) {
    if (nv == 1) return (c >= 0 && c < 4) ? &kSpmv[c] : nullptr;
    return (nv > 1 && c >= 0 && c < 10) ? &kSpmm[c] : nullptr;
}

dspmm_status_t rs_status(rocsparse_status s) {
    switch (s) {
    case rocsparse_status_success: return DSPMM_STATUS_SUCCESS;
    case rocsparse_status_not_implemented:
    case rocsparse_status_invalid_value:
    case rocsparse_status_invalid_size: return DSPMM_STATUS_NOT_SUPPORTED;
    case rocsparse_status_memory_error: return DSPMM_STATUS_ALLOC_FAILED;
    case rocsparse_status_invalid_pointer:
    case rocsparse_status_invalid_handle: return DSPMM_STATUS_INVALID_VALUE;
    default: return DSPMM_STATUS_HIP_ERROR;
    }
}

#define RS(call)                                                   \
    do {                                                           \
        rocsparse_status s_ = (call);                              \
        if (s_ != rocsparse_status_success) return rs_status(s_);  \
    } while (0)

struct Prepared {
    rocsparse_dnvec_descr x = nullptr, y = nullptr;
    rocsparse_dnmat_descr B = nullptr, C = nullptr;
    void *work = nullptr;
    size_t work_bytes = 0;
    bool ready = false;
};

}  // namespace

struct dspmm_rocsparse_s {
    int nr = 0, nc = 0, nnz = 0;
    const int *rowIdx = nullptr;
    const float *dnInput = nullptr;
    float *dnOutput = nullptr;
    hipStream_t stream = nullptr;
    rocsparse_handle handle = nullptr;
    rocsparse_spmat_descr csr = nullptr, coo = nullptr;
    std::map<std::pair<int, int>, Prepared> prepared;
    float one = 1.0f, zero = 0.0f;

    dspmm_status_t call(const Column &col, Prepared &p, int stage, size_t *bytes) {
        rocsparse_spmat_descr A = col.csr ? csr : coo;
        if (p.x)
            return rs_status(rocsparse_spmv(handle, rocsparse_operation_none, &one, A, p.x, &zero,
                                            p.y, rocsparse_datatype_f32_r,
                                            (rocsparse_spmv_alg)col.alg,
                                            (rocsparse_spmv_stage)stage, bytes, p.work));
        return rs_status(rocsparse_spmm(handle, rocsparse_operation_none,
                                        rocsparse_operation_none, &one, A, p.B, &zero, p.C,
                                        rocsparse_datatype_f32_r, (rocsparse_spmm_alg)col.alg,
                                        (rocsparse_spmm_stage)stage, bytes, p.work));
    }

    // Descriptors, workspace and the preprocess stage for (nv, column).
    dspmm_status_t prepare(int nv, int c, Prepared **out) {
        const Column *col = column_of(nv, c);
        if (!col) return DSPMM_STATUS_INVALID_VALUE;
        if (!col->csr && !coo) return DSPMM_STATUS_NOT_SUPPORTED;
        Prepared &p = prepared[{nv, c}];
        if (!p.ready) {
            if (nv == 1) {
                if (!p.x) {
                    RS(rocsparse_create_dnvec_descr(&p.x, nc, (void *)dnInput,
                                                    rocsparse_datatype_f32_r));
                    RS(rocsparse_create_dnvec_descr(&p.y, nr, dnOutput, rocsparse_datatype_f32_r));
                }
            } else if (!p.B) {
                const bool rm = col->layout == DENSE_ROW_MAJOR;
                const rocsparse_order ord = rm ? rocsparse_order_row : rocsparse_order_column;
                RS(rocsparse_create_dnmat_descr(&p.B, nc, nv, rm ? nv : nc, (void *)dnInput,
                                                rocsparse_datatype_f32_r, ord));
                RS(rocsparse_create_dnmat_descr(&p.C, nr, nv, rm ? nv : nr, dnOutput,
                                                rocsparse_datatype_f32_r, ord));
            }
            size_t bytes = 0;
            dspmm_status_t st = call(*col, p, 1 /* buffer_size */, &bytes);
            if (st) return st;
            if (bytes > p.work_bytes) {
                if (p.work) (void)hipFree(p.work);
                p.work = nullptr;
                p.work_bytes = 0;
                if (hipMalloc(&p.work, bytes) != hipSuccess) return DSPMM_STATUS_ALLOC_FAILED;
                p.work_bytes = bytes;
            }
            st = call(*col, p, 2 /* preprocess */, &bytes);
            if (st) return st;
            p.ready = true;
        }
        *out = &p;
        return DSPMM_STATUS_SUCCESS;
    }

    ~dspmm_rocsparse_s() {
        for (auto &kv : prepared) {
            Prepared &p = kv.second;
            if (p.x) rocsparse_destroy_dnvec_descr(p.x);
            if (p.y) rocsparse_destroy_dnvec_descr(p.y);
            if (p.B) rocsparse_destroy_dnmat_descr(p.B);
            if (p.C) rocsparse_destroy_dnmat_descr(p.C);
            if (p.work) (void)hipFree(p.work);
        }
        if (csr) rocsparse_destroy_spmat_descr(csr);
        if (coo) rocsparse_destroy_spmat_descr(coo);
        if (handle) rocsparse_destroy_handle(handle);
    }
};

extern "C" {

dspmm_status_t dspmm_rocsparse_create(dspmm_rocsparse_t *h, int nr, int nc, int nnz,
                                      const int *rowPtr, const int *rowIdx, const int *colIdx,
                                      const float *values, const float *dnInput, float *dnOutput,
                                      dspmm_stream_t stream) {
    if (!h || nr < 0 || nc < 0 || nnz < 0 || !rowPtr || !colIdx || !values || !dnInput ||
        !dnOutput)
        return DSPMM_STATUS_INVALID_VALUE;
    *h = nullptr;
    auto *r = new (std::nothrow) dspmm_rocsparse_s();
    if (!r) return DSPMM_STATUS_ALLOC_FAILED;
    r->nr = nr;
    r->nc = nc;
    r->nnz = nnz;
    r->rowIdx = rowIdx;
    r->dnInput = dnInput;
    r->dnOutput = dnOutput;
    r->stream = reinterpret_cast<hipStream_t>(stream);
    rocsparse_status s = rocsparse_create_handle(&r->handle);
    if (s == rocsparse_status_success) s = rocsparse_set_stream(r->handle, r->stream);
    if (s == rocsparse_status_success)
        s = rocsparse_create_csr_descr(&r->csr, nr, nc, nnz, (void *)rowPtr, (void *)colIdx,
                                       (void *)values, rocsparse_indextype_i32,
                                       rocsparse_indextype_i32, rocsparse_index_base_zero,
                                       rocsparse_datatype_f32_r);
    if (s == rocsparse_status_success && rowIdx)
        s = rocsparse_create_coo_descr(&r->coo, nr, nc, nnz, (void *)rowIdx, (void *)colIdx,
                                       (void *)values, rocsparse_indextype_i32,
                                       rocsparse_index_base_zero, rocsparse_datatype_f32_r);
    if (s != rocsparse_status_success) {
        delete r;
        return rs_status(s);
    }
    *h = r;
    return DSPMM_STATUS_SUCCESS;
}

dspmm_status_t dspmm_rocsparse_destroy(dspmm_rocsparse_t h) {
    delete h;
    return DSPMM_STATUS_SUCCESS;
}

int dspmm_rocsparse_columns(int nv) { return nv == 1 ? 4 : (nv > 1 ? 10 : 0); }

const char *dspmm_rocsparse_column_name(int nv, int column) {
    const Column *c = column_of(nv, column);
    return c ? c->name : "";
}

int dspmm_rocsparse_column_layout(int nv, int column) {
    const Column *c = column_of(nv, column);
    return c ? c->layout : -1;
}

dspmm_status_t dspmm_rocsparse_run(dspmm_rocsparse_t h, int nv, int column) {
    if (!h) return DSPMM_STATUS_INVALID_VALUE;
    Prepared *p = nullptr;
    dspmm_status_t st = h->prepare(nv, column, &p);
    if (st) return st;
    size_t bytes = p->work_bytes;
    return h->call(*column_of(nv, column), *p, 3 /* compute */, &bytes);
}

dspmm_status_t dspmm_rocsparse_benchmark(dspmm_rocsparse_t h, int nv, int column,
                                         int warmupIters, int repeatIters, float *ms) {
    if (!h || !ms || warmupIters < 0 || repeatIters < 1) return DSPMM_STATUS_INVALID_VALUE;
    *ms = -1.0f;
    for (int i = 0; i < warmupIters; i++) {
        dspmm_status_t st = dspmm_rocsparse_run(h, nv, column);
        if (st) return st;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return DSPMM_STATUS_HIP_ERROR;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return DSPMM_STATUS_HIP_ERROR;
    }
    dspmm_status_t st = dspmm_rocsparse_run(h, nv, column);  // prepares outside the timing
    if (!st && hipEventRecord(e0, h->stream) != hipSuccess) st = DSPMM_STATUS_HIP_ERROR;
    for (int i = 0; !st && i < repeatIters; i++) st = dspmm_rocsparse_run(h, nv, column);
    if (!st && hipEventRecord(e1, h->stream) != hipSuccess) st = DSPMM_STATUS_HIP_ERROR;
    if (!st && hipEventSynchronize(e1) != hipSuccess) st = DSPMM_STATUS_HIP_ERROR;
    float t = 0;
    if (!st && hipEventElapsedTime(&t, e0, e1) != hipSuccess) st = DSPMM_STATUS_HIP_ERROR;
    if (!st) *ms = t / repeatIters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

}  // extern "C"
