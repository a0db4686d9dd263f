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
#include "pbe_device.hpp"
#include "pbe_internal.hpp"

namespace pbe {

struct AggArgs {
    const double* __restrict__ N;
    const double* __restrict__ x;
    double* rhs;
    int n;
    int nb;              // tile blocks per axis
    KernelCoef coef;
    double log_x0;
    double inv_log_r;
};

template <int U>
struct AggShape {
    static constexpr int T = kWave * U;   // tile edge (rows == columns)
    static constexpr int RW = T / 4;      // rows per wave
    static constexpr int EXTRA = 256;     // near-birth reach beyond the tile
    static constexpr int WIN = T + EXTRA; // LDS window of target bins
};

// Tiles before diagonal d in diagonal-major order.
__device__ __forceinline__ long long tiles_before(long long d, long long nb)
{
    return d * nb - d * (d - 1) / 2;
}

template <int WIN>
struct BirthSink {
    double* win;
    double* rhs;
    int k0;
    __device__ __forceinline__ void put(int i, double c) const
    {
        const int o = i - k0;
        if (o >= 0 && o < WIN) lds_add(&win[o], c);
        else global_add(&rhs[i], c);
    }
};

// Birth of one near pair: clip rules of src/aggregation_kernels.cuh:191-195
// (upper clip tested first), otherwise the exact bracket and the K-R split
// of :196-200 and :227-231.
template <int WIN>
__device__ __forceinline__ void near_birth(const AggArgs& a, const BirthSink<WIN>& s,
                                        double v, double rate, int k,
                                        double xfirst, double xlast)
{
    const int n = a.n;
    if (v >= xlast) {
        s.put(n - 1, rate);
    } else if (v <= xfirst) {
        s.put(0, rate);
    } else {
        const int hi = (a.inv_log_r != 0.0)
            ? bracket_geometric(a.x, n, v, a.log_x0, a.inv_log_r)
            : bracket_search(a.x, k, n - 1, v);
        const int lo = hi - 1;
        const double xl = a.x[lo], xh = a.x[hi];
        const double w = (v - xl) / (xh - xl);
        s.put(lo, (1.0 - w) * rate);
        s.put(hi, w * rate);
    }
}

template <int KT, int U>
__global__ __launch_bounds__(256)
void agg_tile_kernel(AggArgs a)
{
    using K = Kern<KT>;
    using S = AggShape<U>;
    constexpr int R = K::R;
    constexpr int RS = (R + 2) & ~1;      // row record: R factors + x_j, 16-B aligned
    constexpr int T = S::T, RW = S::RW, WIN = S::WIN;

    __shared__ double rowbuf[4][RW * RS];
    __shared__ double colwin[WIN];
    __shared__ double red[4][R];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid >> 6;
    const int n = a.n;
    const long long nb = a.nb;

    // ---- tile decode (diagonal-major) ----
    const long long t = blockIdx.x;
    const double Bq = 2.0 * (double)nb + 1.0;
    long long d = (long long)floor(0.5 * (Bq - sqrt(Bq * Bq - 8.0 * (double)t)));
    if (d < 0) d = 0;
    if (d > nb - 1) d = nb - 1;
    while (d > 0 && tiles_before(d, nb) > t) --d;
    while (d + 1 < nb && tiles_before(d + 1, nb) <= t) ++d;
    const int p = (int)(t - tiles_before(d, nb));
    const int j0 = p * T;
    const int k0 = (p + (int)d) * T;

    for (int i = tid; i < WIN; i += 256) colwin[i] = 0.0;

    // ---- row factors of this wave's rows -> LDS ----
    const int rw0 = j0 + wv * RW;
    const int rw1 = min(rw0 + RW, n);
    double* myrows = rowbuf[wv];
    double tj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tj[r] = 0.0;
    for (int rr = lane; rr < RW; rr += kWave) {
        const int j = rw0 + rr;
        double av[R];
        double xj = 0.0;
        if (j < n) {
            xj = a.x[j];
            const double Nj = a.N[j];
            double cb, icb;
            pivot_roots<KT>(xj, cb, icb);
            K::row(av, xj, cb, icb, a.coef);
#pragma unroll
            for (int r = 0; r < R; ++r) av[r] *= Nj;
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) av[r] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            myrows[rr * RS + r] = av[r];
            tj[r] += av[r];
        }
        myrows[rr * RS + R] = xj;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) tj[r] = wave_sum(tj[r]);

    // ---- column factors of this lane's columns ----
    double Bc[U][R], xk[U], xk1[U], idx[U];
    int kc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = k0 + lane + kWave * u;
        kc[u] = k;
        if (k < n) {
            const double xv = a.x[k];
            const double Nk = a.N[k];
            double cb, icb, f[R];
            pivot_roots<KT>(xv, cb, icb);
            K::col(f, xv, cb, icb);
#pragma unroll
            for (int r = 0; r < R; ++r) Bc[u][r] = Nk * f[r];
            xk[u] = xv;
            if (k < n - 1) {
                const double x1 = a.x[k + 1];
                xk1[u] = x1;
                idx[u] = 1.0 / (x1 - xv);
            } else {
                xk1[u] = xv;     // never "far": every pair of column n-1 clips
                idx[u] = 0.0;
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) Bc[u][r] = 0.0;
            xk[u] = 0.0;
            xk1[u] = 0.0;
            idx[u] = 0.0;
        }
    }
    const double xfirst = a.x[0];
    const double xlast = a.x[n - 1];
    __syncthreads();

    // ---- death of this row block (diagonal tiles only) ----
    if (d == 0) {
        double s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = 0.0;
        for (int k = tid; k < n; k += 256) {
            const double xv = a.x[k];
            const double Nk = a.N[k];
            double cb, icb, f[R];
            pivot_roots<KT>(xv, cb, icb);
            K::col(f, xv, cb, icb);
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = fma(Nk, f[r], s[r]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            s[r] = wave_sum(s[r]);
            if (lane == 0) red[wv][r] = s[r];
        }
        __syncthreads();
        double Sg[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Sg[r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
        for (int rr = lane; rr < RW; rr += kWave) {
            const int j = rw0 + rr;
            if (j < n) {
                double dth = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) dth = fma(myrows[rr * RS + r], Sg[r], dth);
                lds_add(&colwin[j - k0], -dth);
            }
        }
    }

    // ---- births of this wave's pairs ----
    if (rw0 < rw1) {
        const int jmax = rw1 - 1;
        const double xjmax = a.x[jmax];
        bool ok = true;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kc[u];
            ok = ok && (k >= n || (k <= n - 2 && jmax < k && (xjmax + xk[u]) < xk1[u]));
        }
        const int nrows = rw1 - rw0;
        if (__all(ok)) {
            // every pair lands in (k, k+1): registers only
            double accw[U], colsum[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                double cs = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) cs = fma(tj[r], Bc[u][r], cs);
                colsum[u] = cs;
                accw[u] = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) Bc[u][r] *= idx[u];
            }
            for (int jj = 0; jj < nrows; ++jj) {
                const double* rp = myrows + jj * RS;
                double av[R];
#pragma unroll
                for (int r = 0; r < R; ++r) av[r] = rp[r];
                const double xj = rp[R];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    double pr = av[0] * Bc[u][0];
#pragma unroll
                    for (int r = 1; r < R; ++r) pr = fma(av[r], Bc[u][r], pr);
                    const double v = xj + xk[u];
                    const double dv = v - xk[u];
                    accw[u] = fma(dv, pr, accw[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = kc[u];
                if (k < n) {
                    lds_add(&colwin[k - k0], colsum[u] - accw[u]);
                    lds_add(&colwin[k + 1 - k0], accw[u]);
                }
            }
        } else {
            const BirthSink<WIN> sink{colwin, a.rhs, k0};
            double accr[U], accw[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { accr[u] = 0.0; accw[u] = 0.0; }
            for (int jj = 0; jj < nrows; ++jj) {
                const int j = rw0 + jj;
                const double* rp = myrows + jj * RS;
                double av[R];
#pragma unroll
                for (int r = 0; r < R; ++r) av[r] = rp[r];
                const double xj = rp[R];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = kc[u];
                    if (k < n && j <= k) {
                        double pr = av[0] * Bc[u][0];
#pragma unroll
                        for (int r = 1; r < R; ++r) pr = fma(av[r], Bc[u][r], pr);
                        if (j == k) pr *= 0.5;
                        const double v = xj + xk[u];
                        if (v < xk1[u]) {
                            accr[u] += pr;
                            accw[u] = fma((v - xk[u]) * idx[u], pr, accw[u]);
                        } else {
                            near_birth<WIN>(a, sink, v, pr, k, xfirst, xlast);
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = kc[u];
                if (k < n - 1) {
                    lds_add(&colwin[k - k0], accr[u] - accw[u]);
                    lds_add(&colwin[k + 1 - k0], accw[u]);
                }
            }
        }
    }
    __syncthreads();

    // ---- flush the window ----
    for (int i = tid; i < WIN; i += 256) {
        const int k = k0 + i;
        const double v = colwin[i];
        if (k < n && v != 0.0) global_add(&a.rhs[k], v);
    }
}

template <int KT, int U>
static hipError_t launch_agg_t(const AggArgs& args, hipStream_t stream)
{
    const long long nb = args.nb;
    const long long tiles = nb * (nb + 1) / 2;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((agg_tile_kernel<KT, U>), dim3((unsigned)tiles), dim3(256),
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
