// agg_kernels.hip -- fixed-pivot aggregation source term on gfx950.
//
// Reference semantics: src/aggregation_kernels.cuh:145-239 (per pair (j <= k):
// rate = beta N_j N_k, halved on the diagonal; death of rate at j and k; birth
// of rate split by Kumar-Ramkrishna weights between the pivots bracketing
// v = x_j + x_k, clipped to bin n-1 / bin 0 outside the grid).
//
// Design (MI355X-first, not the reference's one-thread-per-pair layout):
//
//  * Square pair tiles of T = 64*U rows x T columns over the upper triangle,
//    one 256-thread workgroup (4 wave64) per tile, tiles dispatched
//    diagonal-major so the costlier diagonal-band tiles start first.  Lanes
//    own columns k (U each), waves own a quarter of the rows; row factors
//    are staged in LDS and read as wave-uniform broadcasts.
//
//  * The pair rate uses the separable form of the collision kernel
//    (pbe_device.hpp): R <= 5 multiply-adds per pair.
//
//  * "Far" pairs -- v = x_j + x_k < x_{k+1} -- are the vast majority at large
//    n; all of them land in the bracket (k, k+1), so a lane keeps that
//    column's births in registers: sum_j w rate (to k+1) and the lo part
//    sum_j (1-w) rate = colsum_k - sum_j w rate, where colsum_k is the
//    separable column sum of the rates.  A wave whose rows are all far from
//    its columns runs a 3-op-plus-rate inner loop with no compares.
//
//  * "Near" pairs (diagonal band, clipped pairs) take the exact bracket
//    (log-ratio guess plus fix-up on geometric grids, binary search
//    otherwise) and ds_add_f64 into an LDS window of the tile's target bins.
//
//  * Death N_i sum_k beta_ik N_k is separable too; the diagonal tile of each
//    row block adds it for its rows from R global column sums.
//
//  * One flush per tile: the LDS window is added to rhs with FP64 global
//    atomics (zero entries skipped).
#include "agg_tile.hpp"

namespace pbe {

template <int KT, int U>
__global__ __launch_bounds__(256)
void agg_tile_kernel(AggArgs a)
{
    __shared__ AggSmem<KT, U> sm;
    constexpr int T = AggShape<U>::T, S = AggShape<U>::SPLIT, RW = AggShape<U>::RW;
    const long long nb = a.nb;
    const long long bt = band_tiles(nb);
    const long long t = blockIdx.x;
    int p, d;
    if (t < S * bt) {
        const long long tt = t / S;
        const int q = (int)(t - tt * S);
        d = tt < nb ? 0 : 1;
        p = (int)(tt < nb ? tt : tt - nb);
        agg_tile_body<KT, U>(a, p * T, (p + d) * T, d == 0, p * T + q * (T / S), RW / S, sm);
    } else {
        decode_tile(t - S * bt + bt, nb, p, d);
        agg_tile_body<KT, U>(a, p * T, (p + d) * T, d == 0, p * T, RW, sm);
    }
}

template <int KT, int U>
static hipError_t launch_agg_t(const AggArgs& args, hipStream_t stream)
{
    const long long nb = args.nb;
    const long long tiles = nb * (nb + 1) / 2;
    const long long items = tiles + (AggShape<U>::SPLIT - 1) * band_tiles(nb);
    if (items > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((agg_tile_kernel<KT, U>), dim3((unsigned)items), dim3(256),
                       0, stream, args);
    return hipSuccess;
}

template <int U>
static hipError_t launch_agg_u(const AggArgs& args, int kernel, hipStream_t stream)
{
    switch (kernel) {
    case 0: return launch_agg_t<0, U>(args, stream);
    case 1: return launch_agg_t<1, U>(args, stream);
    case 2: return launch_agg_t<2, U>(args, stream);
    case 3: return launch_agg_t<3, U>(args, stream);
    case 4: return launch_agg_t<4, U>(args, stream);
    case 5: return launch_agg_t<5, U>(args, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_aggregation(const double* N, const double* x, double* rhs,
                              const pbe_aggregation_params& p, int tile_u,
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
    switch (tile_u) {
    case 1:
        args.nb = (p.n + AggShape<1>::T - 1) / AggShape<1>::T;
        return launch_agg_u<1>(args, p.kernel_type, stream);
    case 2:
        args.nb = (p.n + AggShape<2>::T - 1) / AggShape<2>::T;
        return launch_agg_u<2>(args, p.kernel_type, stream);
    case 8:
        args.nb = (p.n + AggShape<8>::T - 1) / AggShape<8>::T;
        return launch_agg_u<8>(args, p.kernel_type, stream);
    default:
        args.nb = (p.n + AggShape<4>::T - 1) / AggShape<4>::T;
        return launch_agg_u<4>(args, p.kernel_type, stream);
    }
}

}  // namespace pbe
