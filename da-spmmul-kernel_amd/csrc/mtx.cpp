// mtx.cpp -- Matrix-Market coordinate loader (host) and the sorted-COO ->
// CSR row-offset build, restating the reference's loader path:
//   readMtx           util/spmatrix.hpp:286-333
//   mm_read_banner    util/mmio.hpp:215-298  (case-insensitive tokens)
//   mm_read_mtx_crd_size util/mmio.hpp:308-336 (skips '%' comment lines)
//   readTuples        util/spmatrix.hpp:104-216 (1-based -> 0-based,
//                     pattern -> 1.0, integer and real values as float)
//   makeSymmetric     util/spmatrix.hpp:218-284 (symmetric files only)
//   customSort        util/spmatrix.hpp:75-102 (row, then column)
//
// By default the output equals the reference's for the same file, including
// two behaviours of makeSymmetric that a caller may not want:
//   * diagonal entries of symmetric files are dropped (remove_self_loops
//     defaults to true, :223,253-254) -- DSPMM_MTX_KEEP_DIAGONAL keeps them;
//   * when removed entries are compacted away, row and column indices move
//     but values do not (:264-283), so values after the first removed entry
//     belong to other entries -- DSPMM_MTX_ALIGN_VALUES moves them too.
// Entries with equal (row, column) keep file order (the reference's
// std::sort leaves their order unspecified).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.hpp"

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *path) : f(std::fopen(path, "r")) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

enum Field { F_REAL, F_INTEGER, F_PATTERN, F_COMPLEX };
enum Sym { S_GENERAL, S_SYMMETRIC, S_HERMITIAN, S_SKEW };

void lower(char *p) {
    for (; *p; p++) *p = (char)std::tolower((unsigned char)*p);
}

// util/mmio.hpp:215-298
bool read_banner(FILE *f, Field &field, Sym &sym) {
    char line[1025], banner[64], mtx[64], crd[64], dt[64], ss[64];
    if (!std::fgets(line, sizeof line, f)) return false;
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dt, ss) != 5) return false;
    lower(mtx);
    lower(crd);
    lower(dt);
    lower(ss);
    if (std::strncmp(banner, "%%MatrixMarket", 14) != 0) return false;
    if (std::strcmp(mtx, "matrix") != 0) return false;
    if (std::strcmp(crd, "coordinate") != 0) return false;   // dense arrays: not a sparse file
    if (!std::strcmp(dt, "real")) field = F_REAL;
    else if (!std::strcmp(dt, "integer")) field = F_INTEGER;
    else if (!std::strcmp(dt, "pattern")) field = F_PATTERN;
    else if (!std::strcmp(dt, "complex")) field = F_COMPLEX;
    else return false;
    if (!std::strcmp(ss, "general")) sym = S_GENERAL;
    else if (!std::strcmp(ss, "symmetric")) sym = S_SYMMETRIC;
    else if (!std::strcmp(ss, "hermitian")) sym = S_HERMITIAN;
    else if (!std::strcmp(ss, "skew-symmetric")) sym = S_SKEW;
    else return false;
    return true;
}

// util/mmio.hpp:308-336
bool read_size(FILE *f, int &M, int &N, int &nz) {
    char line[1025];
    M = N = nz = 0;
    do {
        if (!std::fgets(line, sizeof line, f)) return false;
    } while (line[0] == '%');
    if (std::sscanf(line, "%d %d %d", &M, &N, &nz) == 3) return true;
    // A blank line: keep scanning.  Text that is not an integer stops here
    // (fscanf would return 0 forever without consuming it).
    int got;
    do {
        got = std::fscanf(f, "%d %d %d", &M, &N, &nz);
        if (got == EOF || got == 0) return false;
    } while (got != 3);
    return true;
}

// Stable sort of the three arrays by (row, column).
void sort_rc(std::vector<int> &r, std::vector<int> &c, std::vector<float> &v) {
    std::vector<size_t> idx(r.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b];
    });
    std::vector<int> r2(r.size()), c2(c.size());
    std::vector<float> v2(v.size());
    for (size_t i = 0; i < idx.size(); i++) {
        r2[i] = r[idx[i]];
        c2[i] = c[idx[i]];
        v2[i] = v[idx[i]];
    }
    r.swap(r2);
    c.swap(c2);
    v.swap(v2);
}

// util/spmatrix.hpp:218-284
void make_symmetric(std::vector<int> &r, std::vector<int> &c, std::vector<float> &v,
                    bool remove_diagonal, bool align_values) {
    const size_t n0 = r.size();
    for (size_t i = 0; i < n0; i++) {
        if (c[i] != r[i]) {
            r.push_back(c[i]);
            c.push_back(r[i]);
            v.push_back(v[i]);
        }
    }
    sort_rc(r, c, v);
    const int n = (int)r.size();
    if (n == 0) return;
    int curr = c[0], curr_row = r[0];
    for (int i = 0; i < n; i++) {
        const int last = curr, last_row = curr_row;
        curr = c[(size_t)i];
        curr_row = r[(size_t)i];
        if (remove_diagonal && curr_row == curr) c[(size_t)i] = -1;        // :253-254
        if (i > 0 && curr == last && curr_row == last_row) c[(size_t)i] = -1;  // :258
    }
    // :264-277 -- compaction: each removed slot i takes the next kept entry
    // at or after i + shift; when none is left the array ends at i.
    int shift = 0;
    for (int i = 0; i + shift < n; i++) {
        if (c[(size_t)i] != -1) continue;
        int back = i + shift;
        while (back < n && c[(size_t)back] == -1) back++;
        shift = back - i;
        if (back == n) break;
        c[(size_t)i] = c[(size_t)back];
        r[(size_t)i] = r[(size_t)back];
        if (align_values) v[(size_t)i] = v[(size_t)back];
        c[(size_t)back] = -1;
    }
    const size_t keep = (size_t)(n - shift);
    r.resize(keep);
    c.resize(keep);
    v.resize(keep);
}

}  // namespace

extern "C" {

dspmm_status_t dspmm_mtx_read(const char *path, unsigned flags, dspmm_coo *out) {
    if (!path || !out) return DSPMM_STATUS_INVALID_VALUE;
    std::memset(out, 0, sizeof(*out));
    File file(path);
    if (!file.f) return DSPMM_STATUS_INVALID_VALUE;
    Field field;
    Sym sym;
    if (!read_banner(file.f, field, sym)) return DSPMM_STATUS_INVALID_VALUE;
    if (field == F_COMPLEX) return DSPMM_STATUS_NOT_SUPPORTED;
    int M, N, nz;
    if (!read_size(file.f, M, N, nz) || M < 0 || N < 0 || nz < 0) return DSPMM_STATUS_INVALID_VALUE;
    std::vector<int> r, c;
    std::vector<float> v;
    r.reserve((size_t)nz);
    c.reserve((size_t)nz);
    v.reserve((size_t)nz);
    for (int i = 0; i < nz; i++) {                    // util/spmatrix.hpp:123-137, 180-190
        int ri, ci;
        if (std::fscanf(file.f, "%d", &ri) != 1) return DSPMM_STATUS_INVALID_VALUE;
        if (std::fscanf(file.f, "%d", &ci) != 1) return DSPMM_STATUS_INVALID_VALUE;
        float val = 1.0f;
        if (field == F_INTEGER) {
            int iv;
            if (std::fscanf(file.f, "%d", &iv) != 1) return DSPMM_STATUS_INVALID_VALUE;
            val = (float)iv;
        } else if (field == F_REAL) {
            if (std::fscanf(file.f, "%f", &val) != 1) return DSPMM_STATUS_INVALID_VALUE;
        }
        if (ri < 1 || ri > M || ci < 1 || ci > N) return DSPMM_STATUS_INVALID_VALUE;
        r.push_back(ri - 1);
        c.push_back(ci - 1);
        v.push_back(val);
    }
    if (sym == S_SYMMETRIC)
        make_symmetric(r, c, v, !(flags & DSPMM_MTX_KEEP_DIAGONAL), (flags & DSPMM_MTX_ALIGN_VALUES) != 0);
    sort_rc(r, c, v);                                 // util/spmatrix.hpp:327
    const size_t n = r.size();
    out->nr = M;
    out->nc = N;
    out->nnz = (int)n;
    out->rowIdx = (int *)std::malloc(sizeof(int) * std::max<size_t>(n, 1));
    out->colIdx = (int *)std::malloc(sizeof(int) * std::max<size_t>(n, 1));
    out->values = (float *)std::malloc(sizeof(float) * std::max<size_t>(n, 1));
    if (!out->rowIdx || !out->colIdx || !out->values) {
        dspmm_coo_free(out);
        return DSPMM_STATUS_ALLOC_FAILED;
    }
    std::copy(r.begin(), r.end(), out->rowIdx);
    std::copy(c.begin(), c.end(), out->colIdx);
    std::copy(v.begin(), v.end(), out->values);
    return DSPMM_STATUS_SUCCESS;
}

void dspmm_coo_free(dspmm_coo *coo) {
    if (!coo) return;
    std::free(coo->rowIdx);
    std::free(coo->colIdx);
    std::free(coo->values);
    coo->rowIdx = coo->colIdx = nullptr;
    coo->values = nullptr;
}

// Sorted COO rows -> rowPtr[nr+1] (generateCsrRowPtrArray,
// spmvspmm_benchmark.cu:41-76; also defined for nnz == 0).
dspmm_status_t dspmm_csr_rowptr(int nr, int nnz, const int *rowIdx, int *rowPtr) {
    if (nr < 0 || nnz < 0 || !rowPtr || (nnz > 0 && !rowIdx)) return DSPMM_STATUS_INVALID_VALUE;
    for (int i = 0; i <= nr; i++) rowPtr[i] = 0;
    for (int i = 0; i < nnz; i++) {
        if (rowIdx[i] < 0 || rowIdx[i] >= nr || (i && rowIdx[i] < rowIdx[i - 1]))
            return DSPMM_STATUS_INVALID_VALUE;
        rowPtr[rowIdx[i] + 1]++;
    }
    for (int i = 0; i < nr; i++) rowPtr[i + 1] += rowPtr[i];
    return DSPMM_STATUS_SUCCESS;
}

}  // extern "C"
