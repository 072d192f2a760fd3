// spmvspmm -- the sweep harness: loads a Matrix-Market file, checks every
// algorithm against a host loop, then prints one CSV row per dense width
// nv = 1, 2, 4, ..., maxNv with GFLOP/s of the 8 reference algorithms and
// the rocSPARSE comparator columns.
//
// Follows spmvspmm_benchmark.cu:329-419 (main), :155-223 (checkAlgorithms),
// :237-327 (benchmarkAlgorithms), :79-90 (calcMaxNv) and run.sh (the CSV
// header).  Usage, as the reference's binary:
//
//   spmvspmm <file.mtx>                    check, then print CSV rows
//   spmvspmm <file.mtx> profile <maxNv>    one roctx-bracketed launch per
//                                          algorithm and nv (dry_run)
//
// Options (anywhere after the file): --native appends a column with the
// library's own nnz-tiled kernel (row-major, heuristic configuration);
// --seed S seeds the dense input (default 0).  The reference seeds its
// input with time(0) and draws (rand() % 1000 - 500) / 1000; this harness
// draws from the same set of values with a fixed seed.
//
// Exit status: 0 when every check passed, 1 otherwise (the reference returns
// -1), 2 on a usage or load error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "dspmm.h"
#include "dspmm_rocsparse.h"
#include "spmvspmm.hpp"

namespace {

const int kMinNv = 1;
const int kMaxNv = 128;  // spmvspmm_benchmark.cu:5

const SPMV_SPMM_ALG kAlgs[4] = {ALG_CSR_SCALAR, ALG_CSR_VECTOR, ALG_COO_SCALAR, ALG_COO_VECTOR};
const DenseLayout kLayouts[2] = {DENSE_ROW_MAJOR, DENSE_COL_MAJOR};

// calcMaxNv (spmvspmm_benchmark.cu:79-90): the largest power of two <= 128
// whose dense input and output fit beside the sparse matrix.
int calc_max_nv(int nr, int nc, int nnz) {
    dspmm_device_info info;
    dspmm::check(dspmm_get_device_info(&info), "device info");
    const size_t sp = (size_t)(nr + 1 + 2 * (size_t)nnz) * sizeof(int) + (size_t)nnz * sizeof(float);
    const long long fit = (long long)((info.total_mem - sp) / sizeof(float) / (size_t)(nr + nc));
    int nv = kMaxNv;
    while (nv > fit && nv > 0) nv /= 2;
    return nv;
}

// The host product C = A * B in FP32, summing each row in storage order
// (cpuSpmvReference, spmvspmm_benchmark.cu:11-39).
void host_product(DenseLayout layout, int nr, int nc, int nv, const std::vector<int> &rowPtr,
                  const std::vector<int> &colIdx, const std::vector<float> &values,
                  const std::vector<float> &B, std::vector<float> &C) {
    for (int r = 0; r < nr; r++)
        for (int k = 0; k < nv; k++) {
            float acc = 0.0f;
            for (int p = rowPtr[r]; p < rowPtr[r + 1]; p++) {
                const float b = layout == DENSE_ROW_MAJOR ? B[(size_t)colIdx[p] * nv + k]
                                                          : B[(size_t)k * nc + colIdx[p]];
                const float prod = values[p] * b;
                acc = acc + prod;
            }
            if (layout == DENSE_ROW_MAJOR) C[(size_t)r * nv + k] = acc;
            else C[(size_t)k * nr + r] = acc;
        }
}

// checkResult (spmvspmm_benchmark.cu:93-116): absolute tolerance 1e-3, the
// first mismatch is reported.
bool check_result(DenseLayout layout, int nr, int nv, const std::vector<float> &ref,
                  const std::vector<float> &out, const std::string &what) {
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < nv; j++) {
            const size_t idx = layout == DENSE_ROW_MAJOR ? (size_t)i * nv + j : (size_t)j * nr + i;
            if (std::fabs(out[idx] - ref[idx]) > 1e-3f) {
                std::fprintf(stderr, "Check Failed %s nv=%d out[%d,%d] %f != %f\n", what.c_str(),
                             nv, i, j, out[idx], ref[idx]);
                return false;
            }
        }
    return true;
}

struct Args {
    std::string file;
    bool profile = false;
    int profile_nv = 0;
    bool native = false;
    unsigned seed = 0;
};

bool parse(int argc, char **argv, Args &a) {
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        if (s == "--native") a.native = true;
        else if (s == "--seed" && i + 1 < argc) a.seed = (unsigned)std::strtoul(argv[++i], 0, 10);
        else pos.push_back(s);
    }
    if (pos.empty()) return false;
    a.file = pos[0];
    if (pos.size() > 1 && pos[1] == "profile") {
        a.profile = true;
        a.profile_nv = pos.size() > 2 ? std::atoi(pos[2].c_str()) : 0;
    }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    Args args;
    if (!parse(argc, argv, args)) {
        std::cerr << "usage: spmvspmm <file.mtx> [profile <maxNv>] [--native] [--seed S]"
                  << std::endl;
        return 2;
    }
    try {
        dspmm_coo coo;
        dspmm_status_t st = dspmm_mtx_read(args.file.c_str(), 0, &coo);
        if (st != DSPMM_STATUS_SUCCESS) {
            std::cerr << "Could not read " << args.file << ": " << dspmm_get_error_string(st)
                      << std::endl;
            return 2;
        }
        const int nr = coo.nr, nc = coo.nc, nnz = coo.nnz;
        std::vector<int> rowIdx(coo.rowIdx, coo.rowIdx + nnz), colIdx(coo.colIdx, coo.colIdx + nnz);
        std::vector<float> values(coo.values, coo.values + nnz);
        dspmm_coo_free(&coo);
        std::vector<int> rowPtr(nr + 1);
        dspmm::check(dspmm_csr_rowptr(nr, nnz, rowIdx.data(), rowPtr.data()), "csr_rowptr");

        if (nnz == 0 || nr == 0 || nc == 0) {  // the reference reads rowIdx[0] (:44)
            std::cerr << "Error: empty matrix." << std::endl;
            return 2;
        }
        int maxNv = args.profile ? args.profile_nv : calc_max_nv(nr, nc, nnz);
        std::fprintf(stderr, "Read file finished. nrows = %d, ncols = %d, nnz = %d, maxNv = %d\n",
                     nr, nc, nnz, maxNv);
        if (maxNv <= 0) {
            std::cerr << "Error: get maxNv=0. " << std::endl;
            return 2;
        }

        dspmm::CUArray<int> dRowPtr(rowPtr), dRowIdx(rowIdx), dColIdx(colIdx);
        dspmm::CUArray<float> dValues(values);
        dRowPtr.sync_device();
        dRowIdx.sync_device();
        dColIdx.sync_device();
        dValues.sync_device();
        std::vector<float> B((size_t)nc * maxNv);
        std::mt19937 gen(args.seed);
        for (auto &b : B) b = (float)((int)(gen() % 1000) - 500) / 1000;
        dspmm::CUArray<float> dB(B), dC;
        dC.init_zeros((size_t)nr * maxNv);
        dB.sync_device();
        dC.sync_device();

        SpmvSpmmProblem problem(nr, nc, nnz, maxNv, dRowPtr.get_device_data(),
                                dRowIdx.get_device_data(), dColIdx.get_device_data(),
                                dValues.get_device_data(), dB.get_device_data(),
                                dC.get_device_data());

        if (args.profile) {  // benchmarkAlgorithms(profile = true), :261-275
            for (int nv = kMinNv; nv <= maxNv; nv *= 2)
                for (DenseLayout l : kLayouts)
                    for (SPMV_SPMM_ALG alg : kAlgs) problem.dry_run(l, nv, alg, 5);
            dspmm::check(dspmm_device_synchronize(), "synchronize");
            return 0;
        }

        dspmm_rocsparse_t vendor = nullptr;
        dspmm::check(dspmm_rocsparse_create(&vendor, nr, nc, nnz, dRowPtr.get_device_data(),
                                            dRowIdx.get_device_data(), dColIdx.get_device_data(),
                                            dValues.get_device_data(), dB.get_device_data(),
                                            dC.get_device_data(), nullptr),
                     "rocsparse comparator");
        dspmm_matrix_t native = nullptr;
        if (args.native)
            dspmm::check(dspmm_matrix_create(&native, nr, nc, nnz, rowPtr.data(), nullptr,
                                             colIdx.data(), values.data(), 0),
                         "matrix_create");

        // checkAlgorithms (:155-223)
        bool passed = true;
        std::vector<float> ref((size_t)nr * maxNv), out;
        const size_t cbytes = (size_t)nr * sizeof(float);
        for (int nv = kMinNv; nv <= maxNv; nv *= 2) {
            for (DenseLayout l : kLayouts) {
                ref.assign((size_t)nr * nv, 0.0f);
                host_product(l, nr, nc, nv, rowPtr, colIdx, values, B, ref);
                auto fetch = [&]() {
                    dspmm::check(dspmm_device_synchronize(), "synchronize");
                    out.assign((size_t)nr * nv, 0.0f);
                    dspmm::check(dspmm_memcpy_d2h(out.data(), dC.get_device_data(), cbytes * nv),
                                 "memcpy_d2h");
                };
                const int ncols = dspmm_rocsparse_columns(nv);
                for (int c = 0; c < ncols; c++) {
                    if (dspmm_rocsparse_column_layout(nv, c) != (nv == 1 ? DENSE_ROW_MAJOR : l)) continue;
                    dspmm::check(dspmm_fill_zero(dC.get_device_data(), cbytes * nv, nullptr), "zero");
                    st = dspmm_rocsparse_run(vendor, nv, c);
                    if (st == DSPMM_STATUS_NOT_SUPPORTED) continue;
                    dspmm::check(st, "rocsparse run");
                    fetch();
                    passed = check_result(l, nr, nv, ref, out,
                                          std::string("rocsparse_") + dspmm_rocsparse_column_name(nv, c)) && passed;
                }
                for (SPMV_SPMM_ALG alg : kAlgs) {
                    dspmm::check(dspmm_fill_zero(dC.get_device_data(), cbytes * nv, nullptr), "zero");
                    problem.run(l, nv, alg);
                    fetch();
                    passed = check_result(l, nr, nv, ref, out, "alg " + std::to_string((int)alg)) && passed;
                }
                if (native && l == DENSE_ROW_MAJOR) {
                    dspmm::check(dspmm_spmm_ex(native, DENSE_ROW_MAJOR, nv, dB.get_device_data(), nv,
                                               DSPMM_R_32F, dC.get_device_data(), nv, nullptr,
                                               nullptr, 0, nullptr),
                                 "spmm_ex");
                    fetch();
                    passed = check_result(l, nr, nv, ref, out, "native") && passed;
                }
            }
        }
        if (!passed) {
            dspmm_rocsparse_destroy(vendor);
            if (native) dspmm_matrix_destroy(native);
            return 1;
        }

        // benchmarkAlgorithms (:277-326): GFLOP/s = (2 nnz nv / 1e6) / ms
        for (int nv = kMinNv; nv <= maxNv; nv *= 2) {
            const float mflop = (float)nnz / 1e6f * 2 * nv;
            std::cout << args.file << "," << nv;
            for (DenseLayout l : kLayouts)
                for (SPMV_SPMM_ALG alg : kAlgs) std::cout << "," << mflop / problem.benchmark(l, nv, alg);
            for (int c = 0; c < dspmm_rocsparse_columns(nv); c++) {
                float ms = -1;
                st = dspmm_rocsparse_benchmark(vendor, nv, c, 20, 200, &ms);
                if (st == DSPMM_STATUS_NOT_SUPPORTED) std::cout << ",nan";
                else {
                    dspmm::check(st, "rocsparse benchmark");
                    std::cout << "," << mflop / ms;
                }
            }
            if (native) {
                // keep the column aligned with the header on the short nv == 1 rows
                for (int c = dspmm_rocsparse_columns(nv); c < dspmm_rocsparse_columns(2); c++)
                    std::cout << ",";
                float ms = 0;
                dspmm::check(dspmm_benchmark_ex(native, DENSE_ROW_MAJOR, nv, dB.get_device_data(),
                                                nv, DSPMM_R_32F, dC.get_device_data(), nv, nullptr,
                                                5, 50, &ms, nullptr, nullptr, nullptr),
                             "benchmark_ex");
                std::cout << "," << mflop / ms;
            }
            std::cout << std::endl;
        }
        dspmm_rocsparse_destroy(vendor);
        if (native) dspmm_matrix_destroy(native);
        return 0;
    } catch (const dspmm::Error &e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
}
