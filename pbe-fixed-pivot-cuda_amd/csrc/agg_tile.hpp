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
    static constexpr int SPLIT = 4;       // band tiles run as SPLIT row slices
    static constexpr int EXTRA = 256;     // near-birth reach beyond the tile
    static constexpr int WIN = T + EXTRA; // LDS window of target bins
};

// Work items of the single-instance launch: the 2nb - 1 tiles of the
// diagonal band (d = 0, 1), whose near pairs cost several times a far pair,
// first and split into SPLIT row slices each, then the far tiles whole, in
// diagonal-major order.
__host__ __device__ __forceinline__ long long band_tiles(long long nb)
{
    return nb >= 2 ? 2 * nb - 1 : nb;
}

// Tiles before diagonal d in diagonal-major order.
__host__ __device__ __forceinline__ long long tiles_before(long long d, long long nb)
{
    return d * nb - d * (d - 1) / 2;
}

// Tiles of the first D diagonals (d = 0 .. D-1) of an nb-block triangle.
__host__ __device__ __forceinline__ long long band_tiles_d(long long nb, int D)
{
    return tiles_before(D < nb ? D : nb, nb);
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

constexpr int kCursorRows = 4;   // rows whose rates the cursor path forms together

template <int KT, int U>
struct AggSmem {
    static constexpr int R = Kern<KT>::R;
    static constexpr int RS = (R + 2) & ~1;   // row record: R factors + x_j, 16-B aligned
    double rowbuf[4][(AggShape<U>::RW + kCursorRows) * RS];
    double colwin[AggShape<U>::WIN];
    double xw[AggShape<U>::WIN + 1];    // pivots of the window bins
    double idw[AggShape<U>::WIN];       // 1/(x_{i+1} - x_i) of the window bins
    double coln[AggShape<U>::T];        // populations of the tile's columns
    double rowwin[AggShape<U>::T];      // row deaths of an off-diagonal tile
    double colc[AggShape<U>::T];        // cube roots of the tile's column pivots
    double colic[AggShape<U>::T];       // and their reciprocals
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

// Births of a tile's rows by direct brackets, for grids with the geometric
// lookup (inv_log_r > 0).  On a geometric grid of ratio r the event volume
// v = x_j + x_k = x_k (1 + r^-e), e = k - j, lies delta(e) =
// floor(log_r(1 + r^-e)) bins above k whatever k is, so a pair's bracket is
// k + delta(e): no search, and no dependence on the pair before it (the
// cursor path below moves a bracket row by row, one LDS round trip after
// another).  The guess is formed in FP32 with the hardware exp2 / log2 (a
// few 1e-6 bins of error) and checked against the pivots; a pair whose guess
// misses (a rounding edge, a grid that is not quite geometric) or falls
// outside the LDS window takes the exact search.  Rows go kDirectSlots / U
// at a time, their rates, guesses and pivot reads issued together.  Far
// pairs (v < x_{k+1}) sum into registers as in the far path; near pairs add
// their two parts to the LDS window (global atomics beyond it); pairs at or
// past x_{n-1} clip to bin n-1 (src/aggregation_kernels.cuh:191-192).
constexpr int kDirectSlots = 4;

template <int KT, int U>
__device__ __forceinline__ void agg_tile_direct_births(
    const AggArgs& a, const int k0, const int rw0, const int nrows, const double xlast,
    const int (&kc)[U], const double (&xk)[U], const double (&xk1)[U], const double (&idx)[U],
    const double (&Bc)[U][Kern<KT>::R], const double* myrows, AggSmem<KT, U>& sm)
{
    constexpr int R = Kern<KT>::R;
    constexpr int RS = AggSmem<KT, U>::RS;
    constexpr int WIN = AggShape<U>::WIN;
    constexpr int G = kDirectSlots / U > 0 ? kDirectSlots / U : 1;   // rows per group
    const int n = a.n;
    const BirthSink<WIN> sink{sm.colwin, a.rhs, k0};
    const float il2r = (float)a.inv_log_r * 0.693147180559945f;   // 1 / log2 r
    const float l2r = 1.0f / il2r;                                  // log2 r
    auto xat = [&](int i) { return (i - k0 <= WIN) ? lds_ld(&sm.xw[i - k0]) : glb_ld(&a.x[i]); };
    auto idat = [&](int i) {
        const int o = i - k0;
        return o < WIN ? lds_ld(&sm.idw[o]) : 1.0 / (glb_ld(&a.x[i + 1]) - glb_ld(&a.x[i]));
    };
    double cs[U], hw[U], cl[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        cs[u] = 0.0;
        hw[u] = 0.0;
        cl[u] = 0.0;
    }
    for (int jb = 0; jb < nrows; jb += G) {
        double pr[G][U], v[G][U], xc[G][U], idc[G][U];
        int c[G][U], cls[G][U];   // cls 0: no pair, 1: far, 2: near, 3: clipped
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int jj = jb + g;
            const int j = rw0 + jj;
            const double* rp = myrows + (jj < nrows ? jj : 0) * RS;
            double av[R];
#pragma unroll
            for (int r = 0; r < R; ++r) av[r] = rp[r];
            const double xj = rp[R];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = kc[u];
                double p = av[0] * Bc[u][0];
#pragma unroll
                for (int r = 1; r < R; ++r) p = fma(av[r], Bc[u][r], p);
                pr[g][u] = j == k ? 0.5 * p : p;
                const double vv = xj + xk[u];
                v[g][u] = vv;
                const bool has = jj < nrows && j <= k && k < n;
                cls[g][u] = !has ? 0 : (vv >= xlast ? 3 : (vv < xk1[u] ? 1 : 2));
                const float d = __builtin_amdgcn_logf(
                                    1.0f + __builtin_amdgcn_exp2f(-(float)(k - j) * l2r)) * il2r;
                c[g][u] = max(k + (int)d, k + 1);   // a near pair's bracket is above k
            }
        }
        // the guessed brackets' pivots from the window (index clamped: a
        // guess past the window counts as a miss)
        bool miss = false;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int o = min(c[g][u] - k0, WIN - 1);
                const double x0 = sm.xw[o], x1 = sm.xw[o + 1];
                xc[g][u] = x0;
                idc[g][u] = sm.idw[o];
                const bool hit = c[g][u] - k0 < WIN && x0 <= v[g][u] && v[g][u] < x1;
                miss = miss || (cls[g][u] == 2 && !hit);
            }
        if (__any(miss)) {
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (cls[g][u] != 2) continue;
                    const int o = c[g][u] - k0;
                    if (o < WIN && xc[g][u] <= v[g][u] && v[g][u] < xat(c[g][u] + 1)) continue;
                    // exact bracket: x_{k+1} <= v < x_{n-1}
                    const double vv = v[g][u];
                    int cn;
                    if (vv < sm.xw[WIN]) {
                        cn = k0 + window_bracket_lo(sm.xw, kc[u] + 1 - k0, WIN, vv);
                    } else {
                        cn = bracket_geometric(a.x, n, vv, a.log_x0, a.inv_log_r) - 1;
                    }
                    c[g][u] = cn;
                    xc[g][u] = xat(cn);
                    idc[g][u] = idat(cn);
                }
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double p = pr[g][u];
                if (cls[g][u] == 3) {
                    cl[u] += p;
                } else if (cls[g][u] == 1) {
                    cs[u] += p;
                    hw[u] = fma((v[g][u] - xk[u]) * idx[u], p, hw[u]);
                } else if (cls[g][u] == 2) {
                    const double h = (v[g][u] - xc[g][u]) * idc[g][u] * p;
                    sink.put(c[g][u], p - h);
                    sink.put(c[g][u] + 1, h);
                }
            }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = kc[u];
        if (k >= n) continue;
        if (cs[u] != 0.0 || hw[u] != 0.0) {
            sink.put(k, cs[u] - hw[u]);
            sink.put(k + 1, hw[u]);
        }
        if (cl[u] != 0.0) sink.put(n - 1, cl[u]);
    }
}

// One work item of the pair tile with rows [j0, j0+T) and columns [k0, k0+T)
// of the upper triangle of the instance described by `a`: the rows
// [row_base, row_base + 4*rpw), rpw rows per wave (the whole tile, or one
// slice of a costlier diagonal-band tile).  `diag` marks a diagonal tile.
// All 256 threads call it.
template <int KT, int U>
__device__ __forceinline__ void agg_tile_body(const AggArgs& a, const int j0, const int k0,
                                              const bool diag, const int row_base,
                                              const int rpw, AggSmem<KT, U>& sm)
{
    using K = Kern<KT>;
    using S = AggShape<U>;
    constexpr int R = K::R;
    constexpr int RS = AggSmem<KT, U>::RS;
    constexpr int WIN = S::WIN;
    auto& rowbuf = sm.rowbuf;
    double* colwin = sm.colwin;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid >> 6;
    const int n = a.n;

    constexpr int T = AggShape<U>::T;
    constexpr int NW = (WIN + 1 + 255) / 256;   // window pivots per thread
    constexpr int NT = (T + 255) / 256;         // tile columns per thread
    const int rw0 = row_base + wv * rpw;
    const int rw1 = min(rw0 + rpw, n);

    // ---- one load phase: window pivots, column populations, first rows ----
    double xv[NW], xn[NW], nv[NT];
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        const int i = tid + 256 * t, b = k0 + i;
        xv[t] = (i <= WIN && b < n) ? a.x[b] : __builtin_inf();
        xn[t] = (i < WIN && b + 1 < n) ? a.x[b + 1] : __builtin_inf();
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int i = tid + 256 * t, k = k0 + i;
        nv[t] = (i < T && k < n) ? a.N[k] : 0.0;
    }
    double xr0 = 0.0, nr0 = 0.0;
    if (lane < rpw && rw0 + lane < rw1) {
        xr0 = a.x[rw0 + lane];
        nr0 = a.N[rw0 + lane];
    }
    const double xlast = a.x[n - 1];

    for (int i = tid; i < WIN; i += 256) colwin[i] = 0.0;
    for (int i = tid; i < T; i += 256) sm.rowwin[i] = 0.0;
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        const int i = tid + 256 * t;
        if (i <= WIN) sm.xw[i] = xv[t];
        // 1/(x_{i+1} - x_i) of every window bin (the near pairs' weights)
        if (i < WIN) sm.idw[i] = k0 + i < n - 1 ? 1.0 / (xn[t] - xv[t]) : 0.0;
    }
    // column tables: populations, cube roots
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int i = tid + 256 * t, k = k0 + i;
        if (i < T) {
            const double xk0 = xv[t];
            sm.coln[i] = nv[t];
            double cb = 0.0, icb = 0.0;
            if (k < n) pivot_roots<KT>(xk0, cb, icb);
            sm.colc[i] = cb;
            sm.colic[i] = icb;
        }
    }

    // ---- row factors of this wave's rows -> LDS ----
    double* myrows = rowbuf[wv];
    double tj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tj[r] = 0.0;
    for (int rr = lane; rr < rpw; rr += kWave) {
        const int j = rw0 + rr;
        double av[R];
        double xj = 0.0;
        if (j < rw1) {
            xj = rr == lane ? xr0 : a.x[j];
            const double Nj = rr == lane ? nr0 : a.N[j];
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

    __syncthreads();

    // ---- column factors of this lane's columns (from the shared tables) ----
    double Bc[U][R], xk[U], xk1[U], idx[U];
    int kc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int c = lane + kWave * u;
        const int k = k0 + c;
        kc[u] = k;
        if (k < n) {
            const double xv = sm.xw[c];
            const double Nk = sm.coln[c];
            double f[R];
            K::col(f, xv, sm.colc[c], sm.colic[c]);
#pragma unroll
            for (int r = 0; r < R; ++r) Bc[u][r] = Nk * f[r];
            xk[u] = xv;
            if (k < n - 1) {
                xk1[u] = sm.xw[c + 1];
                idx[u] = sm.idw[c];
            } else {
                xk1[u] = xv;     // every pair of column n-1 clips to bin n-1
                idx[u] = 0.0;    // (weight 0 in the far loop)
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) Bc[u][r] = 0.0;
            xk[u] = 0.0;
            xk1[u] = 0.0;
            idx[u] = 0.0;
        }
    }

    // ---- death, separable within the tile ----
    // Pairs (j < k) of this tile remove sum_k rate_jk from row j and
    // sum_j rate_jk from column k; with the tile's column sums
    // TK_r = sum_k B_r(k) and row sums TJ_r = sum_j A_r(j) these are
    // sum_r A_r(j) TK_r and sum_r B_r(k) TJ_r.  On a diagonal tile (rows ==
    // columns) bin i loses sum_{k in block} beta_ik N_i N_k = sum_r A_r(i) TK_r,
    // which also covers the diagonal pair (2 x rate/2).  Summed over all tiles
    // this is the full death N_i sum_k beta_ik N_k of the reference.
    // Each wave takes the column deaths of its own rows (sum_r B_r(k) tj_r,
    // tj_r over the wave's rows), so no wave waits for another's sums.
    {
        double tk[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double sacc = 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u) sacc += Bc[u][r];
            tk[r] = wave_sum(sacc);
        }
        for (int rr = lane; rr < rpw; rr += kWave) {
            const int j = rw0 + rr;
            if (j < n) {
                double dth = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) dth = fma(myrows[rr * RS + r], tk[r], dth);
                if (diag) lds_add(&colwin[j - k0], -dth);
                else sm.rowwin[j - j0] -= dth;   // rows of this wave only: no race
            }
        }
        if (!diag && rw0 < rw1) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = kc[u];
                if (k < n) {
                    double dth = 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) dth = fma(Bc[u][r], tj[r], dth);
                    lds_add(&colwin[k - k0], -dth);
                }
            }
        }
    }

    // ---- births of this wave's pairs ----
    if (rw0 < rw1) {
        const int jmax = rw1 - 1;
        const double xjmax = myrows[(jmax - rw0) * RS + R];   // this wave's last row
        // every pair far (v < x_{k+1}), or in column n-1, whose pairs all
        // clip to bin n-1 (src/aggregation_kernels.cuh:191-192): there the
        // far loop's weight factor 1/(x_{k+1} - x_k) is 0, so the column's
        // whole rate goes to bin n-1
        bool ok = true;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kc[u];
            ok = ok && (k >= n || (jmax < k && (k == n - 1 || (xjmax + xk[u]) < xk1[u])));
        }
        const int nrows = rw1 - rw0;
        if (__all(ok)) {
            // every pair lands in (k, k+1) or clips: registers only
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
        } else if (a.inv_log_r > 0.0) {
            agg_tile_direct_births<KT, U>(a, k0, rw0, nrows, xlast, kc, xk, xk1, idx, Bc,
                                          myrows, sm);
        } else {
            // Cursor sweep.  Rows run upward, so for each column the event
            // volume v = x_j + x_k and its exact bracket (cur, cur+1) only
            // move up: a lane keeps the bracket and the running sums
            // sum rate and sum w*rate of the pairs landing in it, and emits
            // them when the bracket moves on.  Far pairs are the case cur = k.
            const BirthSink<WIN> sink{colwin, a.rhs, k0};
            auto xat = [&](int i) { return (i - k0 <= WIN) ? lds_ld(&sm.xw[i - k0]) : glb_ld(&a.x[i]); };
            auto idat = [&](int i) {
                const int o = i - k0;
                return o < WIN ? lds_ld(&sm.idw[o]) : 1.0 / (glb_ld(&a.x[i + 1]) - glb_ld(&a.x[i]));
            };
            // Bracket of v = x_j + x_k beyond (k, k+1): a binary search in the
            // LDS copy of the window's pivots, or over all pivots past it.
            auto jump = [&](int c, double v) {
                if (v < sm.xw[WIN]) return k0 + window_bracket_lo(sm.xw, c - k0, WIN, v);
                const int hi = (a.inv_log_r != 0.0)
                    ? bracket_geometric(a.x, n, v, a.log_x0, a.inv_log_r)
                    : bracket_search(a.x, c, n - 1, v);
                return hi - 1;
            };
            // Each column's bracket at the wave's first row, all columns
            // together (their searches are independent).
            int cur0[U];
            const double xfirst = myrows[R];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = kc[u];
                cur0[u] = k;
                const double v0 = xfirst + xk[u];
                if (k < n && rw0 <= k && v0 < xlast && v0 >= xk1[u]) cur0[u] = jump(k + 1, v0);
            }
            // One column at a time (u outer) keeps a single column's cursor
            // state live, so this path needs few more registers than the
            // far path above.
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = kc[u];
                if (k >= n) continue;
                int cur = cur0[u];
                double xc = cur == k ? xk[u] : xat(cur);
                double xc1 = cur == k ? xk1[u] : xat(cur + 1);
                double idc = cur == k ? idx[u] : idat(cur);
                double accr = 0.0, accw = 0.0, accc = 0.0;
                const int jend = min(nrows, k - rw0 + 1);   // rows j <= k
                // Rates of four rows at a time first (four independent FP64
                // chains), then the bracket logic for them in row order.
                for (int jb = 0; jb < jend; jb += kCursorRows) {
                double prb[kCursorRows], vb[kCursorRows];
#pragma unroll
                for (int q = 0; q < kCursorRows; ++q) {
                    const double* rp = myrows + (jb + q) * RS;   // rows past nrows hold zeros
                    double p = rp[0] * Bc[u][0];
#pragma unroll
                    for (int r = 1; r < R; ++r) p = fma(rp[r], Bc[u][r], p);
                    prb[q] = (rw0 + jb + q == k) ? 0.5 * p : p;
                    vb[q] = rp[R] + xk[u];
                }
#pragma unroll
                for (int q = 0; q < kCursorRows; ++q) {
                    if (jb + q >= jend) break;
                    const double pr = prb[q];
                    const double v = vb[q];
                    if (v >= xlast) {
                        // upper clip (src/aggregation_kernels.cuh:191-192)
                        accc += pr;
                    } else {
                        if (v >= xc1) {
                            if (accr != 0.0 || accw != 0.0) {
                                sink.put(cur, accr - accw);
                                sink.put(cur + 1, accw);
                            }
                            int c = cur + 1;
                            if (xat(c + 1) <= v) c = jump(c, v);
                            cur = c;
                            xc = xat(c);
                            xc1 = xat(c + 1);
                            idc = idat(c);
                            accr = 0.0;
                            accw = 0.0;
                        }
                        accr += pr;
                        accw = fma((v - xc) * idc, pr, accw);
                    }
                }
                }
                if (accr != 0.0 || accw != 0.0) {
                    sink.put(cur, accr - accw);
                    sink.put(cur + 1, accw);
                }
                if (accc != 0.0) sink.put(n - 1, accc);
            }
        }
    }
    __syncthreads();

    // ---- flush the windows ----
    for (int i = tid; i < WIN; i += 256) {
        const int k = k0 + i;
        const double v = colwin[i];
        if (k < n && v != 0.0) global_add(&a.rhs[k], v);
    }
    if (!diag) {
        for (int i = tid; i < AggShape<U>::T; i += 256) {
            const int j = j0 + i;
            const double v = sm.rowwin[i];
            if (j < n && v != 0.0) global_add(&a.rhs[j], v);
        }
    }
}


// Diagonal band: the 2 nb - 1 tiles with d = 0 (diagonal) and d = 1 (the
// block below), each split into S row slices; item t -> tile t / S, slice
// t % S, band tiles in order p = 0.. for d = 0, then for d = 1.
template <int KT, int U>
__device__ __forceinline__ void band_item(const AggArgs& a, int t, int S, AggSmem<KT, U>& sm)
{
    constexpr int T = AggShape<U>::T, RW = AggShape<U>::RW;
    const int tt = t / S, sl = t - tt * S;
    const int d = tt < a.nb ? 0 : 1;
    const int p = tt < a.nb ? tt : tt - a.nb;
    agg_tile_body<KT, U>(a, p * T, (p + d) * T, d == 0, p * T + sl * (T / S), RW / S, sm);
}

// Item t of the single tile launch: the tiles of the first D diagonals
// (those holding near pairs: D = 2 on grids whose near-pair reach is within
// a tile) as S row slices each, first, then the other tiles whole, all in
// diagonal-major order.
template <int KT, int U>
__device__ __forceinline__ void agg_tiles_item(const AggArgs& a, long long t, int S, int D,
                                               AggSmem<KT, U>& sm)
{
    constexpr int T = AggShape<U>::T, RW = AggShape<U>::RW;
    const long long bt = band_tiles_d(a.nb, D);
    int p, d;
    if (t < S * bt) {
        const long long tt = t / S;
        const int sl = (int)(t - tt * S);
        decode_tile(tt, a.nb, p, d);
        agg_tile_body<KT, U>(a, p * T, (p + d) * T, d == 0, p * T + sl * (T / S), RW / S, sm);
    } else {
        decode_tile(t - S * bt + bt, a.nb, p, d);
        agg_tile_body<KT, U>(a, p * T, (p + d) * T, d == 0, p * T, RW, sm);
    }
}

}  // namespace pbe
