// agg_kernels.hip -- fixed-pivot aggregation source term on gfx950.
//
// Reference semantics: src/aggregation_kernels.cuh:145-239 (per pair (j <= k):
// rate = beta N_j N_k, halved on the diagonal; death of rate at j and k; birth
// of rate split by Kumar-Ramkrishna weights between the pivots bracketing
// v = x_j + x_k, clipped to bin n-1 / bin 0 outside the grid).
//
// Design (MI355X-first, not the reference's one-thread-per-pair layout):
//
//  * Column strips of T = 64*U columns, cut into chunks of 64 rows over the
//    upper triangle and split by an estimated cost over exactly the number
//    of workgroups the device holds at once (agg_strip.hpp).  Lanes own
//    columns k (U each), waves own chunks of rows; row factors are staged in
//    LDS and read as wave-uniform broadcasts.
//
//  * The pair rate uses the separable form of the collision kernel
//    (pbe_device.hpp): R <= 5 multiply-adds per pair.
//
//  * "Far" pairs -- v = x_j + x_k < x_{k+1} -- are the vast majority at large
//    n; all of them land in the bracket (k, k+1), so a lane keeps that
//    column's births in registers: sum_j w rate (to k+1) and the lo part
//    sum_j (1-w) rate = colsum_k - sum_j w rate, where colsum_k is the
//    separable column sum of the rates.  A chunk whose rows are all far from
//    the wave's columns runs a 3-op-plus-rate inner loop with no compares.
//    Column n-1 joins them: all its pairs clip to bin n-1 (weight w = 0).
//
//  * "Near" pairs (diagonal band) keep a monotone exact bracket per column
//    (a cursor) and ds_add_f64 the per-bracket sums into an LDS window of the
//    block's target bins (global atomics beyond it).
//
//  * Death N_i sum_k beta_ik N_k is separable: a row chunk below the block
//    loses sum_r A_r(j) TK_r (TK_r: the block's column sums) and the block's
//    columns lose sum_r B_r(k) TJ_r; rows inside the block lose
//    sum_r A_r(i) TK_r, which covers both sides of the block's own pairs.
//
//  * One flush per visit of a block: the LDS window is added to rhs with
//    FP64 global atomics (zero entries skipped).
#include <atomic>

#include "agg_strip.hpp"

namespace pbe {

// At least 3 waves per SIMD (3 workgroups per CU, VGPRs capped at 168), 2
// for the widest strips.
template <int KT, int U>
__global__ __launch_bounds__(256, (U >= 8 ? 2 : 3))
void agg_strip_kernel(AggArgs a, StripPlan plan)
{
    __shared__ StripSmem<KT, U> sm;
    const long long C = plan.total, G = gridDim.x, g = blockIdx.x;
    int qa, ca, qb, cb;
    strip_first_unit(plan, C * g / G, qa, ca);
    strip_first_unit(plan, C * (g + 1) / G, qb, cb);
    for (int q = qa, c = ca; q < qb || (q == qb && c < cb); ++q, c = 0) {
        const int hi = (q == qb) ? cb : strip_nc(plan, q);
        if (c < hi) agg_strip_visit<KT, U>(a, q, c, hi, sm);
    }
}

// Workgroups of one kernel resident on the current device at once.
template <int KT, int U>
static int resident_groups()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&agg_strip_kernel<KT, U>), 256, 0) !=
            hipSuccess || per_cu <= 0)
        per_cu = 1;
    v = cus * per_cu;
    cache[dev].store(v, std::memory_order_relaxed);
    return v;
}

template <int KT, int U>
static hipError_t launch_agg_t(const AggArgs& args, int near_weight, hipStream_t stream)
{
    StripPlan plan;
    plan.U = U;
    plan.W = near_weight;
    plan.nb = (args.n + 64 * U - 1) / (64 * U);
    plan.nch = (args.n + kStripRows - 1) / kStripRows;
    plan.total = strip_total(plan);
    long long G = resident_groups<KT, U>();
    if (G > plan.total) G = plan.total;
    if (G < 1) G = 1;
    hipLaunchKernelGGL((agg_strip_kernel<KT, U>), dim3((unsigned)G), dim3(256), 0, stream,
                       args, plan);
    return hipSuccess;
}

template <int U>
static hipError_t launch_agg_u(const AggArgs& args, int kernel, int w, hipStream_t stream)
{
    switch (kernel) {
    case 0: return launch_agg_t<0, U>(args, w, stream);
    case 1: return launch_agg_t<1, U>(args, w, stream);
    case 2: return launch_agg_t<2, U>(args, w, stream);
    case 3: return launch_agg_t<3, U>(args, w, stream);
    case 4: return launch_agg_t<4, U>(args, w, stream);
    case 5: return launch_agg_t<5, U>(args, w, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_aggregation(const double* N, const double* x, double* rhs,
                              const pbe_aggregation_params& p, int tile_u, int near_weight,
                              hipStream_t stream)
{
    AggArgs args;
    args.N = N;
    args.x = x;
    args.rhs = rhs;
    args.n = p.n;
    args.coef = KernelCoef{p.beta0, p.beta_br, p.beta_sh};
    args.log_x0 = p.log_x0;
    args.inv_log_r = p.inv_log_r;
    args.nb = 0;
    const int w = near_weight < 1 ? 1 : near_weight;
    switch (tile_u) {
    case 1: return launch_agg_u<1>(args, p.kernel_type, w, stream);
    case 2: return launch_agg_u<2>(args, p.kernel_type, w, stream);
    case 8: return launch_agg_u<8>(args, p.kernel_type, w, stream);
    default: return launch_agg_u<4>(args, p.kernel_type, w, stream);
    }
}

}  // namespace pbe
