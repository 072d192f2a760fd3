// Test driver for the reference's Matrix-Market loader.  It is compiled
// against the reference's own headers where they lie (util/spmatrix.hpp,
// util/mmio.hpp; see Makefile.ref) into oracle/_ref/mtx_ref and is run only
// by the test suite, as the checker for dspmm_mtx_read.
//
// Usage: mtx_ref <file.mtx> <out.bin>
// out.bin: int32 nr, nc, nnz; int32 rowIdx[nnz]; int32 colIdx[nnz];
//          float32 values[nnz] -- exactly what readMtx<float>
//          (util/spmatrix.hpp:286-333) returns, as called by
//          spmvspmm_benchmark.cu:349.
#include <cstdio>
#include <vector>

#include "spmatrix.hpp"

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <file.mtx> <out.bin>\n", argv[0]);
        return 2;
    }
    std::vector<int> rows, cols;
    std::vector<float> vals;
    int nr = 0, nc = 0, nnz = 0;
    readMtx<float>(argv[1], rows, cols, vals, nr, nc, nnz);
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 3;
    const int n = (int)rows.size();
    const int head[3] = {nr, nc, n};
    std::fwrite(head, sizeof(int), 3, o);
    std::fwrite(rows.data(), sizeof(int), rows.size(), o);
    std::fwrite(cols.data(), sizeof(int), cols.size(), o);
    std::fwrite(vals.data(), sizeof(float), vals.size(), o);
    std::fclose(o);
    return 0;
}
