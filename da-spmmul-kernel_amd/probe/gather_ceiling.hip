// gather_ceiling -- the bandwidth ceilings that bound sparse x dense
// products on this GPU, measured without any sparse kernel around them:
//   * streaming read of 2 GiB (the HBM rate a kernel can reach);
//   * gathers of whole N-float rows of a dense B (K = 262144 rows) at the
//     column indices of a 3.19M-nonzero uniform-random pattern, random and
//     sorted order, for N = 128 and 32, and over working sets of 4 MB to
//     512 MB (the L2 / last-level-cache / HBM regimes);
//   * 4-byte gathers of x (N = 1) with coalesced index reads.
// All indices are checked on the host against the buffers before launch.
// Build: make -C da-spmmul-kernel_amd bin/gather_ceiling
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <random>
#include <cstring>

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

// Each group of L lanes (VW floats per lane, L*VW = N) gathers rows idx[p] for
// p in its chunk of CH consecutive indices, UNR at a time, and accumulates.
template <int VW, int L, int UNR>
__global__ __launch_bounds__(256) void k_gather(const int* __restrict__ idx, int n, const float* __restrict__ B,
                                                int ldb, int CH, float* __restrict__ out) {
    const int G = 64 / L;
    const int lane = threadIdx.x % 64, g = lane / L, l = lane % L;
    const long gw = (long)(blockIdx.x * 4 + threadIdx.x / 64) * G + g;   // group id
    const long p0 = gw * CH;
    float acc[VW] = {0};
    for (long p = p0; p < p0 + CH && p < n; p += UNR) {
        float b[UNR][VW];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            if (p + u < n && p + u < p0 + CH) {
                const float* src = B + (long)idx[p + u] * ldb + l * VW;
                if constexpr (VW == 4) { float4 t = *(const float4*)src; b[u][0]=t.x; b[u][1]=t.y; b[u][2]=t.z; b[u][3]=t.w; }
                else if constexpr (VW == 2) { float2 t = *(const float2*)src; b[u][0]=t.x; b[u][1]=t.y; }
                else b[u][0] = *src;
            } else {
#pragma unroll
                for (int k = 0; k < VW; k++) b[u][k] = 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; u++)
#pragma unroll
            for (int k = 0; k < VW; k++) acc[k] += b[u][k];
    }
    float s = 0; for (int k = 0; k < VW; k++) s += acc[k];
    if (s == 123.456f) out[gw] = s;     // practically never; defeats DCE
}

// N = 1: a grid-stride loop over the index stream (coalesced), one 4-byte
// gather of x per index.
__global__ __launch_bounds__(256) void k_gather1(const int* __restrict__ idx, int n,
                                                 const float* __restrict__ x, float* __restrict__ out) {
    float acc = 0.f;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) acc += x[idx[i]];
    if (acc == 123.456f) out[0] = acc;
}

__global__ void k_stream(const float4* __restrict__ a, long n4, float* out) {
    float4 acc = make_float4(0,0,0,0);
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n4; i += (long)gridDim.x * 256) {
        float4 t = a[i]; acc.x += t.x; acc.y += t.y; acc.z += t.z; acc.w += t.w;
    }
    if (acc.x + acc.y + acc.z + acc.w == 123.456f) out[0] = acc.x;
}

template <int VW, int L, int UNR>
float time_gather(const int* d_idx, int n, const float* d_B, int ldb, int CH, float* d_out, int reps) {
    const int G = 64 / L;
    const long groups = (n + CH - 1) / CH;
    const long waves = (groups + G - 1) / G;
    const int blocks = (int)((waves + 3) / 4);
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    hipLaunchKernelGGL((k_gather<VW, L, UNR>), dim3(blocks), dim3(256), 0, 0, d_idx, n, d_B, ldb, CH, d_out);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL((k_gather<VW, L, UNR>), dim3(blocks), dim3(256), 0, 0, d_idx, n, d_B, ldb, CH, d_out);
    CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
    float ms; CK(hipEventElapsedTime(&ms, e0, e1));
    return ms / reps;
}

int main() {
    const int K = 262144, N = 128;
    // c2-like column stream: 3.19M uniform random columns (the exact
    // generator does not matter for bandwidth).
    const int n = 3191107;
    std::vector<int> idx(n);
    std::mt19937 rng(2);
    for (int i = 0; i < n; i++) idx[i] = rng() % K;
    std::vector<int> sorted = idx; std::sort(sorted.begin(), sorted.end());
    for (int v : idx) if (v < 0 || v >= K) { printf("bad index\n"); return 1; }

    int *d_idx, *d_sorted; float *d_B, *d_out, *d_big;
    CK(hipMalloc(&d_idx, n * 4)); CK(hipMalloc(&d_sorted, n * 4));
    CK(hipMalloc(&d_B, (size_t)K * N * 4)); CK(hipMalloc(&d_out, (size_t)n * 4));
    const long big = 2L << 30;   // 2 GiB streaming buffer
    CK(hipMalloc(&d_big, big));
    CK(hipMemcpy(d_idx, idx.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(d_sorted, sorted.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemset(d_B, 0, (size_t)K * N * 4)); CK(hipMemset(d_big, 0, big));

    const int reps = 20;
    {   // streaming read bandwidth
        hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_stream, dim3(256 * 16), dim3(256), 0, 0, (const float4*)d_big, big / 16, d_out);
        CK(hipDeviceSynchronize()); CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_stream, dim3(256 * 16), dim3(256), 0, 0, (const float4*)d_big, big / 16, d_out);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1)); ms /= reps;
        printf("stream read 2 GiB: %.1f us  %.0f GB/s\n", ms * 1e3, big / (ms * 1e-3) / 1e9);
    }
    struct R { const char* name; float ms; double bytes; };
    std::vector<R> res;
    auto add = [&](const char* nm, float ms, double bytes) { res.push_back({nm, ms, bytes});
        printf("%-40s %8.1f us  %6.0f GB/s gathered\n", nm, ms * 1e3, bytes / (ms * 1e-3) / 1e9); fflush(stdout); };
    const double g512 = (double)n * 512, g128 = (double)n * 128, g4 = (double)n * 4;
    add("N=128 random VW2 L64 U8 CH64", time_gather<2, 64, 8>(d_idx, n, d_B, 128, 64, d_out, reps), g512);
    add("N=128 random VW4 L32 U8 CH64", time_gather<4, 32, 8>(d_idx, n, d_B, 128, 64, d_out, reps), g512);
    add("N=128 random VW4 L32 U16 CH128", time_gather<4, 32, 16>(d_idx, n, d_B, 128, 128, d_out, reps), g512);
    add("N=128 random VW2 L64 U16 CH256", time_gather<2, 64, 16>(d_idx, n, d_B, 128, 256, d_out, reps), g512);
    add("N=128 sorted VW2 L64 U8 CH64", time_gather<2, 64, 8>(d_sorted, n, d_B, 128, 64, d_out, reps), g512);
    add("N=32 random VW4 L8 U8 CH32", time_gather<4, 8, 8>(d_idx, n, d_B, 32, 32, d_out, reps), g128);
    add("N=32 random VW4 L8 U16 CH64", time_gather<4, 8, 16>(d_idx, n, d_B, 32, 64, d_out, reps), g128);
    add("N=32 sorted VW4 L8 U8 CH32", time_gather<4, 8, 8>(d_sorted, n, d_B, 32, 32, d_out, reps), g128);
    for (int blocks : {1024, 2048, 4096}) {
        hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_gather1, dim3(blocks), dim3(256), 0, 0, d_idx, n, d_B, d_out);
        CK(hipDeviceSynchronize()); CK(hipEventRecord(e0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_gather1, dim3(blocks), dim3(256), 0, 0, d_idx, n, d_B, d_out);
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        char nm[64]; snprintf(nm, sizeof nm, "N=1 random x (1 MB), %d blocks", blocks);
        add(strdup(nm), ms / reps, g4 + (double)n * 4);
    }
    // working-set sweep: row indices folded into the first Kw rows
    {
        const int kws[] = {8192, 32768, 131072, 262144, 524288, 1048576};
        float* d_Bbig; CK(hipMalloc(&d_Bbig, (size_t)1048576 * 128 * 4));
        CK(hipMemset(d_Bbig, 0, (size_t)1048576 * 128 * 4));
        int* d_fold; CK(hipMalloc(&d_fold, n * 4));
        std::vector<int> fold(n);
        for (int kw : kws) {
            for (int i = 0; i < n; i++) fold[i] = (int)(((unsigned long long)idx[i] * 2654435761ull) % kw);
            CK(hipMemcpy(d_fold, fold.data(), n * 4, hipMemcpyHostToDevice));
            char nm[64]; snprintf(nm, sizeof nm, "N=128 random rows<%d (%d MB)", kw, kw / 2048);
            add(strdup(nm), time_gather<2, 64, 8>(d_fold, n, d_Bbig, 128, 64, d_out, reps), g512);
        }
        CK(hipFree(d_Bbig)); CK(hipFree(d_fold));
    }
    add("N=16 of 128 random VW4 L4 U8 CH32", time_gather<4, 4, 8>(d_idx, n, d_B, 128, 32, d_out, reps), (double)n * 64);
    printf("done\n");
    return 0;
}
