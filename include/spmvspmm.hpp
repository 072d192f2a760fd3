// spmvspmm.hpp -- C++ host interface mirroring the reference's problem API
// (spmvspmm.h:13-50, :202-273) and its device-array helpers (helper.h:78-191)
// on top of the C-ABI in dspmm.h.  Header-only; link with -ldspmm.
//
// Differences from the reference, by design:
//   * HIP errors raise dspmm::Error instead of exiting the process
//     (the reference's CUDA_CHECK calls exit(EXIT_FAILURE), helper.h:25-34);
//   * every call takes an optional stream (the reference uses stream 0);
//   * CUArray's copy constructor is deleted (the reference's would double-free).
// Kept as in the reference: a non power-of-two nv makes run()/benchmark()
// print "nv must be power of 2..." to stderr and return (benchmark returns -1),
// and dry_run() also refuses nv > 32.
#ifndef DSPMM_SPMVSPMM_HPP
#define DSPMM_SPMVSPMM_HPP

#include <iostream>
#include <stdexcept>
#include <string>
#include <cstdlib>
#include <ctime>
#include <vector>

#include "dspmm.h"

namespace dspmm {

struct Error : std::runtime_error {
    dspmm_status_t status;
    Error(dspmm_status_t s, const std::string &what)
        : std::runtime_error(what + ": " + dspmm_get_error_string(s)), status(s) {}
};

inline void check(dspmm_status_t s, const char *what) {
    if (s != DSPMM_STATUS_SUCCESS) throw Error(s, what);
}

// hgpu::InitGen (helper.h:42-76): the C library's rand() seeded once with
// srand(seed) (time(0) by default); int values in [-50, 49], float values
// in [-0.5, 0.499] in steps of 1/1000.
template <typename T>
struct InitGen;
template <>
struct InitGen<int> {
    InitGen() : InitGen((unsigned)std::time(nullptr)) {}
    explicit InitGen(unsigned seed) { std::srand(seed); }
    int gen_next() { return std::rand() % 100 - 50; }
};
template <>
struct InitGen<float> {
    InitGen() : InitGen((unsigned)std::time(nullptr)) {}
    explicit InitGen(unsigned seed) { std::srand(seed); }
    float gen_next() { return (float)(std::rand() % 1000 - 500) / 1000; }
};

// hgpu::CUArray (helper.h:78-154): host vector + device buffer.
template <typename T>
class CUArray {
  public:
    CUArray() = default;
    explicit CUArray(const std::vector<T> &data) : host_(data) {}
    CUArray(const CUArray &) = delete;
    CUArray &operator=(const CUArray &) = delete;
    CUArray(CUArray &&o) noexcept : host_(std::move(o.host_)), dev_(o.dev_) { o.dev_ = nullptr; }
    ~CUArray() { free_(); }

    // CUArray::init_random (helper.h:107-116): n values from InitGen<T>.
    void init_random(size_t n) {
        InitGen<T> gen;
        host_.resize(n);
        for (size_t i = 0; i < n; i++) host_[i] = gen.gen_next();
    }
    void init_random(size_t n, unsigned seed) {
        InitGen<T> gen(seed);
        host_.resize(n);
        for (size_t i = 0; i < n; i++) host_[i] = gen.gen_next();
    }
    void init_zeros(size_t n) { host_.assign(n, T(0)); }
    void assign(const std::vector<T> &data) { free_(); host_ = data; }
    // CUArray::sync_device (helper.h:135-140): (re)allocate and upload.
    void sync_device() {
        free_();
        void *p = nullptr;
        check(dspmm_device_alloc(&p, host_.size() * sizeof(T)), "device_alloc");
        dev_ = static_cast<T *>(p);
        check(dspmm_memcpy_h2d(dev_, host_.data(), host_.size() * sizeof(T)), "memcpy_h2d");
    }
    // CUArray::sync_host (helper.h:142-147).
    std::vector<T> &sync_host() {
        check(dspmm_memcpy_d2h(host_.data(), dev_, host_.size() * sizeof(T)), "memcpy_d2h");
        return host_;
    }
    T *get_device_data() { return dev_; }
    T *get_host_data() { return host_.data(); }
    std::vector<T> &get_host_array() { return host_; }
    size_t get_size() const { return host_.size(); }

  private:
    void free_() {
        if (dev_) dspmm_device_free(dev_);
        dev_ = nullptr;
    }
    std::vector<T> host_;
    T *dev_ = nullptr;
};

// hgpu::GpuTimer (helper.h:158-191).
class GpuTimer {
  public:
    explicit GpuTimer(dspmm_stream_t s = nullptr) : s_(s) { check(dspmm_timer_create(&t_, 2), "timer"); }
    ~GpuTimer() { dspmm_timer_destroy(t_); }
    GpuTimer(const GpuTimer &) = delete;
    GpuTimer &operator=(const GpuTimer &) = delete;
    void start() { dspmm_timer_reset(t_); check(dspmm_timer_mark(t_, s_), "timer start"); }
    void stop() { check(dspmm_timer_mark(t_, s_), "timer stop"); }
    float elapsed_msecs() {
        float ms = 0;
        check(dspmm_timer_intervals(t_, &ms, 1), "timer elapsed");
        return ms;
    }

  private:
    dspmm_timer_t t_ = nullptr;
    dspmm_stream_t s_ = nullptr;
};

}  // namespace dspmm

// The reference's problem descriptor: non-owning device pointers.
struct SpmvSpmmProblem {
    int nr, nc, nnz, maxNv;
    int *rowPtr, *rowIdx, *colIdx;
    float *values, *dnInput, *dnOutput;
    dspmm_stream_t stream = nullptr;

    SpmvSpmmProblem(int nr, int nc, int nnz, int maxNv, int *rowPtr, int *rowIdx, int *colIdx,
                    float *values, float *dnInput, float *dnOutput, dspmm_stream_t s = nullptr)
        : nr(nr), nc(nc), nnz(nnz), maxNv(maxNv), rowPtr(rowPtr), rowIdx(rowIdx), colIdx(colIdx),
          values(values), dnInput(dnInput), dnOutput(dnOutput), stream(s) {}

    // spmvspmm.h:202-218
    void run(DenseLayout layout, int nv, SPMV_SPMM_ALG alg) {
        if ((nv & (nv - 1)) != 0) {
            std::cerr << "nv must be power of 2, less than 32. Got nv = " << nv << std::endl;
            return;
        }
        dspmm::check(dspmm_spmm(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx, values,
                                dnInput, dnOutput, stream),
                     "SpmvSpmmProblem::run");
    }

    // spmvspmm.h:220-243: mean milliseconds per launch.
    float benchmark(DenseLayout layout, int nv, SPMV_SPMM_ALG alg, int warmupIters = 5,
                    int repeatIters = 50) {
        if ((nv & (nv - 1)) != 0) {
            std::cerr << "nv must be power of 2, less than 32. Got nv = " << nv << std::endl;
            return -1;
        }
        float ms = 0;
        dspmm::check(dspmm_benchmark_compat(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx,
                                            values, dnInput, dnOutput, warmupIters, repeatIters,
                                            &ms, nullptr, stream),
                     "SpmvSpmmProblem::benchmark");
        return ms;
    }

    // spmvspmm.h:245-273: warm-ups, then one launch inside a profiler window.
    void dry_run(DenseLayout layout, int nv, SPMV_SPMM_ALG alg, int warmupIters = 0) {
        if ((nv & (nv - 1)) != 0 || nv > 32) {
            std::cerr << "nv must be power of 2, less than 32. Got nv = " << nv << std::endl;
            return;
        }
        dspmm::check(dspmm_dry_run_compat(alg, layout, nr, nc, nnz, nv, rowPtr, rowIdx, colIdx,
                                          values, dnInput, dnOutput, warmupIters, stream),
                     "SpmvSpmmProblem::dry_run");
    }
};

#endif  // DSPMM_SPMVSPMM_HPP
