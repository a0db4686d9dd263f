// rhs_fused.hip -- R_agg(N) + R_brk(N) in one launch, for problems whose
// aggregation is a single tile launch (small n); used by the device
// integrator and the adjoint's forward solves.
//
// At these sizes each launch's own work takes a few microseconds and the
// aggregation tiles and the breakage blocks are independent (both only add
// into rhs with FP64 atomics), so one grid holds the aggregation tile items
// followed by the breakage parent blocks, and the two run side by side
// instead of one launch after the other.  The contributions are those of
// the separate launches (agg_kernels.hip, brk_kernels.hip); only the order
// in which the atomics land differs, as it does between any two runs.
//
// Tried and dropped: folding the RK stage update into the same launch, done
// by the workgroup that finishes last (found with device-scope completion
// counters).  Counting and the last workgroup's pass over the bins took
// 4-5 us per stage at n = 1024 against 2 us for a separate stage kernel
// (config-5 forward solve 93 ms against 75 ms).
#include "agg_tile.hpp"
#include "brk_block.hpp"

namespace pbe {

template <int KT, int U>
union RhsSmem {
    AggSmem<KT, U> agg;
    BrkSmem brk;
};

// Items [0, agg_items): aggregation tile items (agg_tiles_item, D leading
// diagonals as S row slices); then breakage blocks of 256 parents (b.n = 0:
// none), selection `sel`.
template <int KT, int U>
__global__ __launch_bounds__(256, 3)
void rhs_fused_kernel(AggArgs a, int S, int D, int agg_items, BrkArgs b, int sel)
{
    __shared__ RhsSmem<KT, U> sm;
    const int t = blockIdx.x;
    if (t < agg_items) {
        agg_tiles_item<KT, U>(a, t, S, D, sm.agg);
        return;
    }
    const int p0 = (t - agg_items) * kBrkParents;
    switch (sel) {
    case 0: brk_block_body<0>(b, p0, sm.brk); break;
    case 1: brk_block_body<1>(b, p0, sm.brk); break;
    case 2: brk_block_body<2>(b, p0, sm.brk); break;
    default: brk_block_body<3>(b, p0, sm.brk); break;
    }
}

template <int U>
static hipError_t launch_rhs_u(int kernel, dim3 grid, const AggArgs& a, int S, int D,
                               int agg_items, const BrkArgs& b, int sel, hipStream_t s)
{
    const dim3 blk(256);
    switch (kernel) {
    case 0: hipLaunchKernelGGL((rhs_fused_kernel<0, U>), grid, blk, 0, s, a, S, D, agg_items, b, sel); break;
    case 1: hipLaunchKernelGGL((rhs_fused_kernel<1, U>), grid, blk, 0, s, a, S, D, agg_items, b, sel); break;
    case 2: hipLaunchKernelGGL((rhs_fused_kernel<2, U>), grid, blk, 0, s, a, S, D, agg_items, b, sel); break;
    case 3: hipLaunchKernelGGL((rhs_fused_kernel<3, U>), grid, blk, 0, s, a, S, D, agg_items, b, sel); break;
    case 4: hipLaunchKernelGGL((rhs_fused_kernel<4, U>), grid, blk, 0, s, a, S, D, agg_items, b, sel); break;
    case 5: hipLaunchKernelGGL((rhs_fused_kernel<5, U>), grid, blk, 0, s, a, S, D, agg_items, b, sel); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rhs_fused(const RhsProblem& p, const double* N, double* rhs,
                            hipStream_t stream)
{
    if (p.batch || !p.agg || !p.brk) return hipErrorNotSupported;
    const pbe_aggregation_params& ap = *p.agg;
    const pbe_breakage_params& bp = *p.brk;
    const TileChoice tc = tile_choice(ap.n, ap.kernel_type, ap.inv_log_r);
    if (tc.u != 1 && tc.u != 2 && tc.u != 4) return hipErrorNotSupported;
    const TilesShape ts = aggregation_tiles_shape(ap, tc.u, tc.w, tc.f);
    if (ts.items <= 0) return hipErrorNotSupported;
    const long long brk_blocks = (bp.n + kBrkParents - 1) / kBrkParents;
    if (ts.items + brk_blocks > 0x7fffffffLL) return hipErrorNotSupported;

    AggArgs a;
    a.N = N;
    a.x = p.x;
    a.rhs = rhs;
    a.n = ap.n;
    a.nb = (ap.n + kWave * tc.u - 1) / (kWave * tc.u);
    a.coef = KernelCoef{ap.beta0, ap.beta_br, ap.beta_sh};
    a.log_x0 = ap.log_x0;
    a.inv_log_r = ap.inv_log_r;
    const BrkArgs b{N, p.x, p.t_q, p.bw_q, rhs, bp.n, bp.n_quad, bp.S0, bp.v_ref, bp.alpha,
                    bp.v_min};
    const dim3 grid((unsigned)(ts.items + brk_blocks));
    const int S = band_split_or_default(tc.w);
    const int items = (int)ts.items;
    switch (tc.u) {
    case 1: return launch_rhs_u<1>(ap.kernel_type, grid, a, S, ts.diags, items, b, bp.selection, stream);
    case 2: return launch_rhs_u<2>(ap.kernel_type, grid, a, S, ts.diags, items, b, bp.selection, stream);
    default: return launch_rhs_u<4>(ap.kernel_type, grid, a, S, ts.diags, items, b, bp.selection, stream);
    }
}

}  // namespace pbe
