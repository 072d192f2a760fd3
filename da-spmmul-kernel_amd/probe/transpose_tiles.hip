// transpose_tiles -- the rate of the out-of-place FP32 transpose that the
// column-major path uses (B into row-major, C back), over LDS tile shapes,
// on the c2 N=128 operand (262144 x 128 and its transpose, 134 MB).  Each
// line gives the time per transpose (HIP events, mean of 20) and the bytes
// read plus written per second.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                              \
        }                                                                          \
    } while (0)

// dst[c * ldd + r] = src[r * lds + c]: TR x TC tiles through LDS.
template <int TR, int TC>
__global__ __launch_bounds__(256) void k_tile(const float *__restrict__ src, long lds, int rows,
                                              int cols, float *__restrict__ dst, long ldd) {
    __shared__ float tile[TR][TC + 1];
    const int tiles_c = (cols + TC - 1) / TC;
    const int tr = (int)blockIdx.x / tiles_c, tc = (int)blockIdx.x % tiles_c;
    const int r0 = tr * TR, c0 = tc * TC;
    const int t = (int)threadIdx.x;
    for (int k = t / TC; k < TR; k += 256 / TC) {
        const int r = r0 + k, c = c0 + t % TC;
        if (r < rows && c < cols) tile[k][t % TC] = src[(long)r * lds + c];
    }
    __syncthreads();
    for (int k = t / TR; k < TC; k += 256 / TR) {
        const int c = c0 + k, r = r0 + t % TR;
        if (r < rows && c < cols) dst[(long)c * ldd + r] = tile[t % TR][k];
    }
}

template <int TR, int TC>
int time_us(const float *s, long lds, int rows, int cols, float *d, long ldd, float *us) {
    const unsigned tiles = (unsigned)(((rows + TR - 1) / TR) * (long)((cols + TC - 1) / TC));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    for (int i = 0; i < 3; i++)
        hipLaunchKernelGGL((k_tile<TR, TC>), dim3(tiles), dim3(256), 0, 0, s, lds, rows, cols, d, ldd);
    CHECK(hipEventRecord(a));
    for (int i = 0; i < 20; i++)
        hipLaunchKernelGGL((k_tile<TR, TC>), dim3(tiles), dim3(256), 0, 0, s, lds, rows, cols, d, ldd);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    *us = ms / 20 * 1e3f;
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    return 0;
}

template <int TR, int TC>
int line(const float *x, float *y, int K, int N) {
    const double gb = 2.0 * sizeof(float) * (double)K * N / 1e9;
    float u1 = 0, u2 = 0;
    if (time_us<TR, TC>(x, K, N, K, y, N, &u1) || time_us<TR, TC>(x, N, K, N, y, K, &u2)) return 1;
    std::printf("tile %3dx%-3d  N x K -> K x N %7.1f us %5.0f GB/s   K x N -> N x K %7.1f us %5.0f GB/s\n",
                TR, TC, u1, gb / (u1 * 1e-6), u2, gb / (u2 * 1e-6));
    return 0;
}

int main() {
    const int K = 262144, N = 128;
    float *x = nullptr, *y = nullptr;
    CHECK(hipMalloc(&x, sizeof(float) * (size_t)K * N));
    CHECK(hipMalloc(&y, sizeof(float) * (size_t)K * N));
    CHECK(hipMemset(x, 0, sizeof(float) * (size_t)K * N));
    if (line<32, 32>(x, y, K, N) || line<32, 64>(x, y, K, N) || line<32, 128>(x, y, K, N) ||
        line<64, 64>(x, y, K, N) || line<64, 32>(x, y, K, N) || line<128, 32>(x, y, K, N) ||
        line<16, 256>(x, y, K, N) || line<256, 16>(x, y, K, N))
        return 1;
    CHECK(hipFree(x));
    CHECK(hipFree(y));
    std::printf("done\n");
    return 0;
}
