// integrate.hip -- explicit time integration on the device, and its hipGraph.
//
// The reference integrates on the host (examples/common/pbe_examples_utils.hpp:
// 391-457): every RK4 stage downloads k, does the axpy on the CPU and uploads
// the stage again -- 8 transfers, 2 allocations and 4 device syncs per step.
// Here a step is 4 right-hand-side evaluations and 4 small stage kernels on
// one stream, with the reference's arithmetic and its evaluation order
//   stage = N + (0.5*dt)*k1 ; N + (0.5*dt)*k2 ; N + dt*k3
//   N    += (dt/6) * (((k1 + 2 k2) + 2 k3) + k4)
// (no FMA contraction, so the rounding is the reference's).  The stage kernel
// that consumes k also zeroes it for the next evaluation, so no memset sits in
// the sequence and the whole step can be captured into a hipGraph.  Small
// problems evaluate the right-hand side in one launch (rhs_fused.hip).
#include "pbe_device.hpp"
#include "pbe_internal.hpp"

namespace pbe {

__global__ void fill_zero_kernel(double* __restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// mode 1: acc = k;          stage = N + (0.5 dt) k   (ysave = N if given)
// mode 2: acc = acc + 2 k;  stage = N + (0.5 dt) k
// mode 3: acc = acc + 2 k;  stage = N + dt k
// mode 4: Nout = N + (dt/6)(acc + k)
// mode 5 (Euler): Nout = N + dt k
// N is read and Nout written (the same array for a step in place); k is
// zeroed after it is read.
__global__ void rk_stage_kernel(const double* N, double* Nout, double* __restrict__ k,
                                double* __restrict__ acc, double* __restrict__ stage,
                                double* __restrict__ ysave, int n, double dt, int mode)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ki = k[i];
    k[i] = 0.0;
    switch (mode) {
    case 1:
        acc[i] = ki;
        stage[i] = __dadd_rn(N[i], __dmul_rn(__dmul_rn(0.5, dt), ki));
        if (ysave) ysave[i] = N[i];
        break;
    case 2:
        acc[i] = __dadd_rn(acc[i], __dmul_rn(2.0, ki));
        stage[i] = __dadd_rn(N[i], __dmul_rn(__dmul_rn(0.5, dt), ki));
        break;
    case 3:
        acc[i] = __dadd_rn(acc[i], __dmul_rn(2.0, ki));
        stage[i] = __dadd_rn(N[i], __dmul_rn(dt, ki));
        break;
    case 4:
        Nout[i] = __dadd_rn(N[i], __dmul_rn(__ddiv_rn(dt, 6.0), __dadd_rn(acc[i], ki)));
        break;
    default:
        Nout[i] = __dadd_rn(N[i], __dmul_rn(dt, ki));
        break;
    }
}

static inline unsigned blocks_for(int n) { return (unsigned)((n + 255) / 256); }

hipError_t enqueue_fill_zero(double* a, int n, hipStream_t s)
{
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks_for(n)), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

// rhs (already zero) += R(Nin)
hipError_t enqueue_rhs(const RhsProblem& p, const double* Nin, double* rhs, hipStream_t s)
{
    hipError_t e = hipSuccess;
    if (p.batch) return enqueue_batched_rhs(p.batch, Nin, p.x, p.t_q, p.bw_q, rhs, s);
    // small problems: aggregation tiles and breakage blocks in one launch
    e = launch_rhs_fused(p, Nin, rhs, s);
    if (e != hipErrorNotSupported) return e;
    e = hipSuccess;
    if (p.agg) {
        const TileChoice t = tile_choice(p.agg->n, p.agg->kernel_type, p.agg->inv_log_r);
        e = launch_aggregation(Nin, p.x, rhs, *p.agg, t.u, t.w, t.f, s);
        if (e != hipSuccess) return e;
    }
    if (p.brk) {
        e = launch_breakage(Nin, p.x, p.t_q, p.bw_q, rhs, *p.brk, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t enqueue_rk4_step_io(const RhsProblem& p, const double* Nin, double* Nout, double dt,
                               double* work, double* const* stages, double* ysave,
                               hipStream_t s)
{
    const int n = p.n;
    double* k = work;
    double* acc = work + n;
    double* st1 = stages ? stages[0] : work + 2 * (size_t)n;
    double* st2 = stages ? stages[1] : st1;
    double* st3 = stages ? stages[2] : st1;
    const dim3 g(blocks_for(n)), b(256);
    double* const none = nullptr;
    hipError_t e;
    if ((e = enqueue_rhs(p, Nin, k, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(rk_stage_kernel, g, b, 0, s, Nin, Nout, k, acc, st1, ysave, n, dt, 1);
    if ((e = enqueue_rhs(p, st1, k, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(rk_stage_kernel, g, b, 0, s, Nin, Nout, k, acc, st2, none, n, dt, 2);
    if ((e = enqueue_rhs(p, st2, k, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(rk_stage_kernel, g, b, 0, s, Nin, Nout, k, acc, st3, none, n, dt, 3);
    if ((e = enqueue_rhs(p, st3, k, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(rk_stage_kernel, g, b, 0, s, Nin, Nout, k, acc, st1, none, n, dt, 4);
    return hipGetLastError();
}

hipError_t enqueue_rk4_step_stages(const RhsProblem& p, double* N, double dt, double* work,
                                   double* const* stages, hipStream_t s)
{
    return enqueue_rk4_step_io(p, N, N, dt, work, stages, nullptr, s);
}

hipError_t enqueue_rk4_step(const RhsProblem& p, double* N, double dt, double* work,
                            hipStream_t s)
{
    return enqueue_rk4_step_stages(p, N, dt, work, nullptr, s);
}

hipError_t enqueue_euler_step(const RhsProblem& p, double* N, double dt, double* work,
                              hipStream_t s)
{
    hipError_t e;
    if ((e = enqueue_rhs(p, N, work, s)) != hipSuccess) return e;
    double* const none = nullptr;
    hipLaunchKernelGGL(rk_stage_kernel, dim3(blocks_for(p.n)), dim3(256), 0, s, N, N, work, none,
                       none, none, p.n, dt, 5);
    return hipGetLastError();
}

}  // namespace pbe
