// agg_tile.hpp -- one aggregation pair tile (device code shared by the
// single-instance and batched kernels); see agg_kernels.hip for the design.
#pragma once

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
struct AggSmem {
    static constexpr int R = Kern<KT>::R;
    static constexpr int RS = (R + 2) & ~1;   // row record: R factors + x_j, 16-B aligned
    double rowbuf[4][AggShape<U>::RW * RS];
    double colwin[AggShape<U>::WIN];
    double red[4][R];
};

// Diagonal-major tile index t -> (position p along diagonal d) for nb blocks.
__device__ __forceinline__ void decode_tile(long long t, long long nb, int& p, int& d)
{
    const double Bq = 2.0 * (double)nb + 1.0;
    long long dd = (long long)floor(0.5 * (Bq - sqrt(Bq * Bq - 8.0 * (double)t)));
    if (dd < 0) dd = 0;
    if (dd > nb - 1) dd = nb - 1;
    while (dd > 0 && tiles_before(dd, nb) > t) --dd;
    while (dd + 1 < nb && tiles_before(dd + 1, nb) <= t) ++dd;
    p = (int)(t - tiles_before(dd, nb));
    d = (int)dd;
}

// One pair tile: rows [j0, j0+T) x columns [k0, k0+T) of the upper triangle of
// the instance described by `a`; `diag` marks the row block's diagonal tile,
// which also adds that row block's death terms.  All 256 threads call it.
template <int KT, int U>
__device__ __forceinline__ void agg_tile_body(const AggArgs& a, const int j0, const int k0,
                                              const bool diag, AggSmem<KT, U>& sm)
{
    using K = Kern<KT>;
    using S = AggShape<U>;
    constexpr int R = K::R;
    constexpr int RS = AggSmem<KT, U>::RS;
    constexpr int RW = S::RW, WIN = S::WIN;
    auto& rowbuf = sm.rowbuf;
    double* colwin = sm.colwin;
    auto& red = sm.red;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid >> 6;
    const int n = a.n;

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
    if (diag) {
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


}  // namespace pbe
