// mfma_f64_probe -- the FP64 matrix-core rate this GPU sustains.
//
// Every wave issues a long run of independent v_mfma_f64_16x16x4_f64 on
// register operands (KT accumulators, no memory traffic but one store per
// thread at the end), over enough workgroups to fill every CU; the rate is
// 2048 flops per MFMA over the event-timed second launch.  The restart GEMM
// (k_ritz_mfma) is set against this measured ceiling beside the published
// 78.6 TFLOP/s peak.  Prints one line per configuration and, last,
//   mfma_f64_tflops <best rate>
//
//   nutrlan_amd/mfma_f64_probe
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double d4 __attribute__((ext_vector_type(4)));

template <int KT>
__global__ __launch_bounds__(256) void k_mfma_f64_rate(double *out, int iters, double seed)
{
    d4 acc[KT];
#pragma unroll
    for (int q = 0; q < KT; ++q) acc[q] = d4{0.0, 0.0, 0.0, 0.0};
    const double a = seed + threadIdx.x * 1e-3, b = seed - threadIdx.x * 1e-3;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < KT; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < KT; ++q) s += acc[q][0] + acc[q][1] + acc[q][2] + acc[q][3];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;  // one element per thread
}

// TFLOP/s of the second of two launches (blocks x 256 threads), or -1.
template <int KT>
static double rate(int blocks, int iters)
{
    double *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double tf = -1.0;
    float ms = 0.f;
    if (hipMalloc((void **)&out, sizeof(double) * (size_t)blocks * 256) != hipSuccess) return -1.0;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        hipLaunchKernelGGL(k_mfma_f64_rate<KT>, dim3(blocks), dim3(256), 0, nullptr, out, iters, 1.0);
        if (hipEventRecord(e0, nullptr) == hipSuccess) {
            hipLaunchKernelGGL(k_mfma_f64_rate<KT>, dim3(blocks), dim3(256), 0, nullptr, out, iters, 1.0);
            if (hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess && hipGetLastError() == hipSuccess &&
                ms > 0.f) {
                const double mfmas = (double)blocks * 4 * iters * KT;  // 4 waves per workgroup
                tf = mfmas * 2048.0 / (ms * 1e-3) / 1e12;
                printf("KT=%d workgroups=%d: %.3f ms, %.1f TFLOP/s\n", KT, blocks, ms, tf);
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return tf;
}

int main()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        fprintf(stderr, "mfma_f64_probe: no GPU\n");
        return 1;
    }
    double best = -1.0, r;
    if ((r = rate<1>(2048, 4096)) > best) best = r;
    if ((r = rate<4>(2048, 1024)) > best) best = r;
    if ((r = rate<4>(4096, 1024)) > best) best = r;
    if ((r = rate<8>(2048, 512)) > best) best = r;
    if (best <= 0.0) return 1;
    printf("mfma_f64_tflops %.2f\n", best);
    return 0;
}
