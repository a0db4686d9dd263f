// atomic_rates.hip -- FP64 atomic-add throughput of the device: the LDS adds
// (ds_add_f64) the aggregation kernels use for their target-bin windows and
// the global adds (global_atomic_add_f64, no return) of their window flushes.
//
//     pbe-fixed-pivot-cuda_amd/bench/atomic_rates
//
// Every case runs 1024 workgroups of 256 threads (four per CU on a 256-CU
// device), each thread issuing `iters` adds; the time of a second, warm
// launch gives lane-adds per second.  Prints one JSON line per case.
//   lds_distinct : lane l of wave w adds to element 64 w + l (the pattern of
//                  neighbouring columns adding to neighbouring bins)
//   lds_same     : every lane of a wave adds to one element (worst case)
//   global_rhs   : lanes add to consecutive doubles of a 16384-double array
//                  (the size of rhs at n = 16384), workgroups offset so the
//                  same bins are hit by many workgroups, as in the flushes
//   global_large : the same over 2^25 doubles (256 MiB): mostly HBM
#include <hip/hip_runtime.h>
#include <cstdio>

namespace {

template <int PATTERN>
__global__ __launch_bounds__(256) void lds_adds(double* out, int iters)
{
    __shared__ double buf[256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    buf[tid] = 0.0;
    __syncthreads();
    const int idx = PATTERN == 0 ? tid : 64 * wv;
    double v = 1.0 + 1e-3 * lane;
    for (int it = 0; it < iters; ++it) {
        __hip_atomic_fetch_add(&buf[idx], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        v += 1e-9;
    }
    __syncthreads();
    out[blockIdx.x * 256 + tid] = buf[tid];
}

__global__ __launch_bounds__(256) void global_adds(double* buf, int n_addr, int iters)
{
    const int tid = threadIdx.x;
    double v = 1.0 + 1e-3 * tid;
    int i = (blockIdx.x * 256 + tid) % n_addr;
    for (int it = 0; it < iters; ++it) {
        __hip_atomic_fetch_add(&buf[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        i += 256;
        if (i >= n_addr) i -= n_addr;
        v += 1e-9;
    }
}

constexpr int kBlocks = 1024;

template <class F>
int time_it(const char* name, double adds, F launch)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    launch();
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    if (hipEventRecord(e0) != hipSuccess) return 1;
    launch();
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    printf("{\"case\": \"%s\", \"lane_adds\": %.0f, \"ms\": %.4f, \"gadds_per_s\": %.1f}\n", name,
           adds, ms, adds / (ms * 1e-3) / 1e9);
    return (hipEventDestroy(e0) == hipSuccess && hipEventDestroy(e1) == hipSuccess) ? 0 : 1;
}

}  // namespace

int main()
{
    const int n_large = 1 << 25, n_rhs = 16384;
    double *out = nullptr, *big = nullptr;
    if (hipMalloc(&out, sizeof(double) * kBlocks * 256) != hipSuccess) return 1;
    if (hipMalloc(&big, sizeof(double) * n_large) != hipSuccess) return 1;
    if (hipMemset(big, 0, sizeof(double) * n_large) != hipSuccess) return 1;
    int bad = 0;
    {
        const int iters = 4096;
        const double adds = (double)iters * kBlocks * 256;
        bad |= time_it("lds_distinct", adds, [&] {
            hipLaunchKernelGGL(lds_adds<0>, dim3(kBlocks), dim3(256), 0, 0, out, iters);
        });
        bad |= time_it("lds_same", adds / 16, [&] {
            hipLaunchKernelGGL(lds_adds<1>, dim3(kBlocks), dim3(256), 0, 0, out, iters / 16);
        });
    }
    {
        const int iters = 512;
        const double adds = (double)iters * kBlocks * 256;
        bad |= time_it("global_rhs", adds, [&] {
            hipLaunchKernelGGL(global_adds, dim3(kBlocks), dim3(256), 0, 0, big, n_rhs, iters);
        });
        bad |= time_it("global_large", adds, [&] {
            hipLaunchKernelGGL(global_adds, dim3(kBlocks), dim3(256), 0, 0, big, n_large, iters);
        });
    }
    return (hipFree(out) == hipSuccess && hipFree(big) == hipSuccess && !bad) ? 0 : 1;
}
