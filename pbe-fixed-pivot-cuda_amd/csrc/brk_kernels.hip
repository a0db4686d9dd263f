// brk_kernels.hip -- fixed-pivot breakage source term on gfx950.
//
// Reference semantics: src/breakage_kernels.cuh:115-133 (death
// rhs[j] -= S(x_j) N_j, skipped unless S_j != 0 and N_j > 2.2e-308) and
// :160-251 (birth: per parent j and quadrature point q, contrib =
// S_j N_j bw_q lands at v = x_j t_q, clipped to bin 0 for v <= x_0 and to bin
// n-1 for v >= x_{n-1}, otherwise split between the left bracket lo and lo+1
// with eta = (v - x_lo)/(x_{lo+1} - x_lo)).
//
// Design: one launch does death and birth.  A 1024-thread workgroup owns P
// consecutive parents; S_j N_j is evaluated once per parent (one pow per
// parent for the power law, not one per quadrature point), the fragments of
// those parents land in the window of bins just below them, which is staged
// in LDS together with the pivots it covers.  Births and deaths ds_add_f64
// into the window; brackets are found by binary search in the LDS copy of x;
// the window is flushed once with FP64 global atomics.  Targets outside the
// window (t_q >= 1, or fragments below its reach) fall back to a search in
// global x and a global atomic.
#include "brk_block.hpp"

namespace pbe {

// 1024 threads per block of 256 parents: a table of m quadrature points
// gives each thread m / 4 birth items (dependent chains of LDS reads and a
// divide) instead of m, and each SIMD four waves to switch between.
constexpr int kBrkThreads = 1024;

template <int SEL>
__global__ __launch_bounds__(kBrkThreads)
void brk_kernel(BrkArgs a)
{
    __shared__ BrkSmem sm;
    brk_block_body<SEL, kBrkThreads>(a, blockIdx.x * kBrkParents, sm);
}

hipError_t launch_breakage(const double* N, const double* x, const double* t_q,
                           const double* bw_q, double* rhs,
                           const pbe_breakage_params& p, hipStream_t stream)
{
    BrkArgs args{N, x, t_q, bw_q, rhs, p.n, p.n_quad, p.S0, p.v_ref, p.alpha, p.v_min};
    const unsigned grid = (unsigned)((p.n + kBrkParents - 1) / kBrkParents);
    switch (p.selection) {
    case 0: hipLaunchKernelGGL(brk_kernel<0>, dim3(grid), dim3(kBrkThreads), 0, stream, args); break;
    case 1: hipLaunchKernelGGL(brk_kernel<1>, dim3(grid), dim3(kBrkThreads), 0, stream, args); break;
    case 2: hipLaunchKernelGGL(brk_kernel<2>, dim3(grid), dim3(kBrkThreads), 0, stream, args); break;
    case 3: hipLaunchKernelGGL(brk_kernel<3>, dim3(grid), dim3(kBrkThreads), 0, stream, args); break;
    default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

}  // namespace pbe
