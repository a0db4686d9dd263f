// fp64_issue.hip -- FP64 FMA issue rate of the device against waves per SIMD
// and independent chains per wave (the ceilings the pair loops are tuned for).
//
//     pbe-fixed-pivot-cuda_amd/bench/fp64_issue
//
// Each thread runs ILP independent chains x_i <- fma(x_i, a, b) for a fixed
// number of iterations; 256 * waves_per_simd workgroups of 256 threads put
// waves_per_simd waves on every SIMD of a 256-CU device.  Prints one line per
// (ILP, waves per SIMD) with the achieved TFLOP/s (an FMA counts 2).
#include <hip/hip_runtime.h>
#include <cstdio>

template <int ILP>
__global__ __launch_bounds__(256) void fma_chains(double* out, int iters, double a, double b)
{
    double x[ILP];
#pragma unroll
    for (int i = 0; i < ILP; ++i) x[i] = threadIdx.x * 1e-3 + i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < ILP; ++i) x[i] = fma(x[i], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < ILP; ++i) s += x[i];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int ILP>
static int run(double* out, int waves_per_simd)
{
    const int blocks = 256 * waves_per_simd;
    const int iters = 16384 * 8 / ILP;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    hipLaunchKernelGGL(fma_chains<ILP>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999999, 1e-7);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    if (hipEventRecord(e0) != hipSuccess) return 1;
    hipLaunchKernelGGL(fma_chains<ILP>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999999, 1e-7);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    const double flops = 2.0 * ILP * (double)iters * blocks * 256;
    printf("{\"ilp\": %d, \"waves_per_simd\": %d, \"fp64_tflops\": %.2f}\n", ILP, waves_per_simd,
           flops / (ms * 1e-3) / 1e12);
    return (hipEventDestroy(e0) == hipSuccess && hipEventDestroy(e1) == hipSuccess) ? 0 : 1;
}

int main()
{
    double* out = nullptr;
    if (hipMalloc(&out, sizeof(double) * 256 * 8 * 256) != hipSuccess) return 1;
    int bad = 0;
    for (int w : {1, 2, 3, 4, 8}) {
        bad |= run<1>(out, w);
        bad |= run<2>(out, w);
        bad |= run<4>(out, w);
        bad |= run<8>(out, w);
        bad |= run<16>(out, w);
    }
    return (hipFree(out) == hipSuccess && !bad) ? 0 : 1;
}
