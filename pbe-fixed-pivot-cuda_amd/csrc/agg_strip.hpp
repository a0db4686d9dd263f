// agg_strip.hpp -- single-instance aggregation, far-pair part: column strips
// split evenly over a resident grid (device code; the launcher and the
// diagonal-band part are in agg_kernels.hip).
//
// Work decomposition.  Column block q holds the T = 64*U columns
// [qT, qT + T).  Its rows j < (q-1)T -- everything below the block's
// diagonal band, which the band kernel takes -- are cut into chunks of 64
// rows.  The chunks are laid out block after block, rows ascending within a
// block; on a grid whose near-pair reach is below T they are all far chunks
// of equal cost.  The grid is exactly the number of workgroups the device
// holds at once, and workgroup g takes chunks [g C / G, (g+1) C / G): every
// workgroup is resident from the start and all finish together, and a
// workgroup stays within one or two column blocks, so the column set-up
// (pivots, cube roots, the LDS window) and the flush of the window are paid
// once per visit of a block rather than once per tile.
//
// One workgroup per compute unit, of StripWaves<U> waves (2 per SIMD).  A
// visit's rows are split evenly over the waves; all waves hold all T columns
// (U per lane) and take their rows 64 at a time, staging the row factors in
// the wave's own LDS rows.  A chunk that is not entirely far (a
// grid finer than the band) takes a per-pair path with LDS atomics.
#pragma once

#include "pbe_device.hpp"
#include "pbe_internal.hpp"
#include "agg_tile.hpp"

namespace pbe {

constexpr int kStripRows = 64;       // rows per chunk
constexpr int kStripPieceRows = 128; // rows a wave stages and sweeps at a time

// Waves per strip workgroup: two per SIMD.  (Three per SIMD, at most 168
// VGPRs each, measured a few percent slower.)
template <int U>
struct StripWaves {
    static constexpr int value = 8;
};
constexpr int kStripExtra = 256;     // near-birth reach of the LDS window beyond the block

// Chunk layout of one launch: block q >= 2 has (q - 1) U chunks.
struct StripPlan {
    long long total;   // C: number of chunks
    int nb;            // column blocks
    int U;             // chunks per block height (T / 64)
};

__host__ __device__ __forceinline__ int strip_nc(const StripPlan& s, int q)
{
    return q >= 2 ? (q - 1) * s.U : 0;
}

// Chunks of the blocks before q.
__host__ __device__ __forceinline__ long long strip_prefix(const StripPlan& s, int q)
{
    if (q <= 2) return 0;
    const long long m = q - 2;
    return (long long)s.U * m * (m + 1) / 2;
}

__host__ __device__ __forceinline__ long long strip_total(const StripPlan& s)
{
    return strip_prefix(s, s.nb);
}

// Workgroup shares.  A visit's column set-up and window flush, during which
// the workgroup's waves wait for one another, cost about five chunks of pairs,
// so the shares split the cost line evenly: chunk t of block
// q sits at cost t + kVisitCost (q - 1), block q's visit charge lying just
// before its chunk 0.  A share that crosses into a new block carries that
// block's charge and so holds one chunk fewer.
constexpr int kVisitCost = 5;

__host__ __device__ __forceinline__ long long strip_cost_start(const StripPlan& s, int q)
{
    return strip_prefix(s, q) + (long long)kVisitCost * (q - 2);
}

__host__ __device__ __forceinline__ long long strip_cost_total(const StripPlan& s)
{
    return strip_cost_start(s, s.nb);
}

// First chunk at or after cost position f (s.total past the end).
__device__ __forceinline__ long long strip_chunk_at_cost(const StripPlan& s, long long f)
{
    if (f >= strip_cost_total(s)) return s.total;
    // q = m + 2 with U m (m + 1) / 2 + kVisitCost m <= f
    const double b = 0.5 * s.U + kVisitCost;
    int m = (int)floor((sqrt(b * b + 2.0 * s.U * (double)f) - b) / s.U);
    if (m < 0) m = 0;
    while (m > 0 && strip_cost_start(s, m + 2) > f) --m;
    while (m + 3 <= s.nb && strip_cost_start(s, m + 3) <= f) ++m;
    const int q = m + 2;
    long long off = f - strip_cost_start(s, q) - kVisitCost;
    if (off < 0) off = 0;
    return strip_prefix(s, q) + off;
}

// Chunk (q, c) at linear index t; (nb, 0) past the end.
__device__ __forceinline__ void strip_first_unit(const StripPlan& s, long long t, int& q, int& c)
{
    if (t >= s.total) {
        q = s.nb;
        c = 0;
        return;
    }
    // q = m + 2 with U m (m + 1) / 2 <= t < U (m + 1)(m + 2) / 2
    int m = (int)floor(0.5 * (sqrt(1.0 + 8.0 * (double)t / s.U) - 1.0));
    if (m < 0) m = 0;
    while (m > 0 && strip_prefix(s, m + 2) > t) --m;
    while (strip_prefix(s, m + 3) <= t) ++m;
    q = m + 2;
    c = (int)(t - strip_prefix(s, q));
}

// Order this wave's LDS writes before its later LDS reads by other lanes.
// LDS instructions of one wave execute in issue order, so only code motion by
// the compiler has to be stopped (no wait on outstanding global atomics).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_wave_barrier();
    __asm__ __volatile__("" ::: "memory");
}

template <int KT, int U, int NW = StripWaves<U>::value>
struct StripSmem {
    static constexpr int R = Kern<KT>::R;
    static constexpr int RS = (R + 2) & ~1;
    static constexpr int T = kWave * U;
    static constexpr int WIN = T + kStripExtra;
    static constexpr int NWV = NW;
    double rowbuf[NWV][(kStripPieceRows + 2) * RS];   // per wave: row factors A_r(j) and x_j
    double colwin[WIN];                  // births/deaths of the window bins
    double xw[WIN + 1];                  // pivots of the window bins
    double idw[T];                       // 1/(x_{i+1} - x_i) of the block's columns
    double colc[T];                      // cube roots of the block's pivots
    double colic[T];                     // and their reciprocals
    double coln[T];                      // populations of the block's columns
    int prog[NWV];                       // per wave: rows of the visit done so far
};

// Column factors B_r(k) = N_k phi_r(k) of this lane's U columns.
template <int KT, int U, int NW>
__device__ __forceinline__ void strip_col_factors(const AggArgs& a, const StripSmem<KT, U, NW>& sm,
                                                  const int k0, const int lane,
                                                  double (&Bc)[U][Kern<KT>::R])
{
    constexpr int R = Kern<KT>::R;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int c = lane + kWave * u;
        const int k = k0 + c;
        if (k < a.n) {
            double f[R];
            Kern<KT>::col(f, sm.xw[c], sm.colc[c], sm.colic[c]);
            const double Nk = sm.coln[c];
#pragma unroll
            for (int r = 0; r < R; ++r) Bc[u][r] = Nk * f[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) Bc[u][r] = 0.0;
        }
    }
}

// Column data of block q held in registers between visits: window pivots,
// populations and right neighbours of the block's columns.
template <int KT, int U, int NWAVES = StripWaves<U>::value>
struct StripCols {
    static constexpr int T = StripSmem<KT, U, NWAVES>::T;
    static constexpr int WIN = StripSmem<KT, U, NWAVES>::WIN;
    static constexpr int NWV = StripSmem<KT, U, NWAVES>::NWV;
    static constexpr int NTH = kWave * NWV;   // threads of the workgroup
    static constexpr int NW = (WIN + 1 + NTH - 1) / NTH;
    static constexpr int NT = (T + NTH - 1) / NTH;
    double xv[NW], nv[NT], xn[NT];   // as loaded: apply valid_* before use

    // Unconditional loads, indices clamped into the arrays: a load under a
    // branch makes the compiler wait for it (and for every vector-memory
    // operation before it, the window's flush atomics included) at once.
    __device__ __forceinline__ void load(const AggArgs& a, int q)
    {
        const int tid = threadIdx.x, n = a.n, k0 = q * T;
#pragma unroll
        for (int t = 0; t < NW; ++t) xv[t] = a.x[min(k0 + tid + NTH * t, n - 1)];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int k = k0 + tid + NTH * t;
            nv[t] = a.N[min(k, n - 1)];
            xn[t] = a.x[min(k + 1, n - 1)];
        }
    }
    // window pivot t of this thread (inf past the window or the grid)
    __device__ __forceinline__ double pivot(int n, int q, int t) const
    {
        const int i = threadIdx.x + NTH * t, b = q * T + i;
        return (i <= WIN && b < n) ? xv[t] : __builtin_inf();
    }
    __device__ __forceinline__ double pop(int n, int q, int t) const
    {
        const int i = threadIdx.x + NTH * t, k = q * T + i;
        return (i < T && k < n) ? nv[t] : 0.0;
    }
    __device__ __forceinline__ double next_pivot(int n, int q, int t) const
    {
        const int i = threadIdx.x + NTH * t, k = q * T + i;
        return (i < T && k + 1 < n) ? xn[t] : __builtin_inf();
    }
};

// One visit of column block q by the workgroup: chunks [c_lo, c_hi).  On
// entry `cols` holds block q's column data and the LDS window is zero; on exit
// `cols` holds block next_q's (if next_q >= 0, loaded before the window's
// global atomics so those loads do not wait behind them) and this thread's
// window bins are flushed and zero again.
template <int KT, int U, int NW = StripWaves<U>::value>
__device__ __forceinline__ void agg_strip_visit(const AggArgs& a, const int q, const int c_lo,
                                                const int c_hi, const int next_q,
                                                StripCols<KT, U, NW>& cols,
                                                StripSmem<KT, U, NW>& sm)
{
    using K = Kern<KT>;
    using Cols = StripCols<KT, U, NW>;
    constexpr int R = K::R;
    constexpr int RS = StripSmem<KT, U, NW>::RS;
    constexpr int T = StripSmem<KT, U, NW>::T;
    constexpr int WIN = StripSmem<KT, U, NW>::WIN;
    constexpr int NTH = Cols::NTH;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
    const int n = a.n;
    const int k0 = q * T;
    const double xlast = a.x[n - 1];

    // ---- column set-up, once per visit ----
    if (tid < Cols::NWV) sm.prog[tid] = 0;
#pragma unroll
    for (int t = 0; t < Cols::NW; ++t) {
        const int i = tid + NTH * t;
        if (i <= WIN) sm.xw[i] = cols.pivot(n, q, t);
    }
    // the block's columns: 1/(x_{k+1} - x_k) (the near path forms the rest of
    // the window's when it needs them), populations, cube roots
#pragma unroll
    for (int t = 0; t < Cols::NT; ++t) {
        const int i = tid + NTH * t, k = k0 + i;
        if (i < T) {
            const double xk_ = cols.pivot(n, q, t);
            sm.idw[i] = k < n - 1 ? 1.0 / (cols.next_pivot(n, q, t) - xk_) : 0.0;
            sm.coln[i] = cols.pop(n, q, t);
            double cb = 0.0, icb = 0.0;
            if (k < n) pivot_roots<KT>(xk_, cb, icb);
            sm.colc[i] = cb;
            sm.colic[i] = icb;
        }
    }
    __syncthreads();

    double xk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xk[u] = sm.xw[lane + kWave * u];
    // TK_r = sum over the block's columns of B_r(k) (the same in every lane)
    double TK[R];
    {
        double Bc[U][R];
        strip_col_factors<KT, U, NW>(a, sm, k0, lane, Bc);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double s = 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u) s += Bc[u][r];
            TK[r] = wave_sum(s);
        }
    }
    // Far births to k+1 before the factors B_r(k) / (x_{k+1} - x_k):
    // acc_r(k) = sum over far rows j of A_r(j) ((x_j + x_k) - x_k).  One
    // accumulator per (column, r) makes a far pair R FMAs plus the two adds
    // of its weight, and gives U*R independent FP64 chains.
    double acc[U][R];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[u][r] = 0.0;
    double* myrows = sm.rowbuf[wv];
    const BirthSink<WIN> sink{sm.colwin, a.rhs, k0};
    const int rmax = min(n, k0 + T);   // rows j <= last column of the block
    wave_lds_sync();

    // The visit's rows are split evenly over the waves; a wave takes its rows
    // in pieces of up to 128, two rows per lane, whose factor computations
    // (cube-root chains) run interleaved.  Every 32 rows a wave posts its
    // progress in LDS and sets its issue priority from its lead over the
    // workgroup's mean (ahead: low, behind: high): the SIMD otherwise favours
    // its oldest wave, which would finish first while the younger one on the
    // same SIMD lags (measured: 20 us against 30 us) and the workgroup waits
    // for it at the end of the visit.  This lane's rows of the next piece are loaded one piece
    // ahead, before the current piece issues its global atomics: a load
    // issued after an atomic would wait for it (vector memory counts retire
    // in order).
    constexpr int PR = kStripPieceRows / kWave;   // rows per lane in a piece
    const int vr0 = c_lo * kStripRows, vr1 = min(c_hi * kStripRows, rmax);
    const int w_lo = vr0 + (int)((long long)(vr1 - vr0) * wv / Cols::NWV);
    const int w_hi = vr0 + (int)((long long)(vr1 - vr0) * (wv + 1) / Cols::NWV);
    // The loads are unconditional (the index clamped into the arrays; rows
    // past the piece are masked where they are used): loads under a branch
    // make the compiler wait for them at once, which undoes the prefetch.
    auto load_rows = [&](int p, double (&xr)[PR], double (&nr)[PR]) {
#pragma unroll
        for (int t = 0; t < PR; ++t) {
            const int j = min(p + lane + kWave * t, n - 1);
            xr[t] = a.x[j];
            nr[t] = a.N[j];
        }
    };
    auto pace = [&](int done) {
        if (lane == 0) sm.prog[wv] = done;
        wave_lds_sync();
        int tot = 0;
#pragma unroll
        for (int w = 0; w < Cols::NWV; ++w) tot += sm.prog[w];
        const int lead = done * Cols::NWV - tot;   // NWV x (own rows - mean rows)
        constexpr int L = 16 * Cols::NWV;
        set_wave_priority(lead > L ? 0 : lead > 0 ? 1 : lead > -L ? 2 : 3);
    };
    double xj_next[PR], Nj_next[PR];
    load_rows(w_lo, xj_next, Nj_next);
    for (int rw0 = w_lo; rw0 < w_hi; rw0 += kStripPieceRows) {
        pace(rw0 - w_lo);
        const int rw1 = min(rw0 + kStripPieceRows, w_hi);
        double xj[PR], Nj[PR];
#pragma unroll
        for (int t = 0; t < PR; ++t) {
            xj[t] = rw0 + lane + kWave * t < rw1 ? xj_next[t] : 0.0;
            Nj[t] = Nj_next[t];
        }
        load_rows(rw0 + kStripPieceRows, xj_next, Nj_next);
        // row factors of this piece (rows past its end: x = 1 for the roots,
        // factors 0)
        double av[PR][R];
#pragma unroll
        for (int t = 0; t < PR; ++t) {
            const bool valid = rw0 + lane + kWave * t < rw1;
            const double xr = valid ? xj[t] : 1.0;
            double cb, icb;
            pivot_roots<KT>(xr, cb, icb);
            K::row(av[t], xr, cb, icb, a.coef);
            const double nr = valid ? Nj[t] : 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) av[t][r] *= nr;
        }
#pragma unroll
        for (int t = 0; t < PR; ++t) {
            const int j = rw0 + lane + kWave * t;
            if (j < rw1) {
                // death of row j against the block's columns (separable),
                // issued after the next piece's loads, so the pair loop
                // below hides its latency
                double dth = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) dth = fma(av[t][r], TK[r], dth);
                global_add(&a.rhs[j], -dth);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) myrows[(lane + kWave * t) * RS + r] = av[t][r];
            myrows[(lane + kWave * t) * RS + R] = xj[t];
        }
        wave_lds_sync();

        const int nrows = rw1 - rw0;
        const int tl = (nrows - 1) / kWave;   // the piece's last row: lane (nrows-1) % 64, row tl
        double xlast_row = xj[0];
#pragma unroll
        for (int t = 1; t < PR; ++t) xlast_row = tl == t ? xj[t] : xlast_row;
        const double xjmax = __shfl(xlast_row, (nrows - 1) % kWave, kWave);
        bool ok = true;   // strip rows lie below the band: never on the diagonal
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + lane + kWave * u;
            ok = ok && (k >= n - 1 || (xjmax + xk[u]) < sm.xw[lane + kWave * u + 1]);
        }
        if (__all(ok)) {
            // every pair lands in (k, k+1), or clips at column n-1 (taken as
            // weight w = 0 there: 1/(x_{k+1} - x_k) is 0 for k = n-1)
            // Rows in pairs, the second row's update of an accumulator U*R
            // instructions after the first's.  The LDS reads of the next
            // pair are issued before the FMAs of this one, alternating two
            // register sets, so their latency hides behind the FMAs (past
            // the last row they read the wave's two spare rows).
            auto ld2 = [&](int jr, double (&p0)[R + 1], double (&p1)[R + 1]) {
                const double* r0 = myrows + jr * RS;
#pragma unroll
                for (int r = 0; r <= R; ++r) {
                    p0[r] = r0[r];
                    p1[r] = r0[RS + r];
                }
            };
            auto fma2 = [&](const double (&p0)[R + 1], const double (&p1)[R + 1]) {
                double w0[U], w1[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    w0[u] = (p0[R] + xk[u]) - xk[u];
                    w1[u] = (p1[R] + xk[u]) - xk[u];
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int u = 0; u < U; ++u) acc[u][r] = fma(p0[r], w0[u], acc[u][r]);
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int u = 0; u < U; ++u) acc[u][r] = fma(p1[r], w1[u], acc[u][r]);
            };
            double pa0[R + 1], pa1[R + 1], pb0[R + 1], pb1[R + 1];
            ld2(0, pa0, pa1);
            int jj = 0;
            for (; jj + 3 < nrows; jj += 4) {
                if ((jj & 31) == 28) pace(rw0 - w_lo + jj);
                ld2(jj + 2, pb0, pb1);
                __builtin_amdgcn_sched_barrier(0);
                fma2(pa0, pa1);
                ld2(jj + 4, pa0, pa1);
                __builtin_amdgcn_sched_barrier(0);
                fma2(pb0, pb1);
            }
            // 0 to 3 rows left; rows jj and jj + 1 are in pa
            if (jj + 1 < nrows) {
                fma2(pa0, pa1);
                jj += 2;
            }
            if (jj < nrows) {
                const double* r0 = myrows + jj * RS;
                const double x0 = r0[R];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double w0 = (x0 + xk[u]) - xk[u];
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[u][r] = fma(r0[r], w0, acc[u][r]);
                }
            }
        } else {
            // Pairs near the diagonal band: each pair adds its two
            // Kumar-Ramkrishna parts to the LDS window (global atomics
            // beyond it).  A lane keeps the exact bracket (cur, cur+1) of
            // v = x_j + x_k for its column; rows run upward, so it only moves
            // up, mostly by at most one bin per row.  One column at a time
            // keeps a single bracket live.
            auto xat = [&](int i) { return (i - k0 <= WIN) ? lds_ld(&sm.xw[i - k0]) : glb_ld(&a.x[i]); };
            auto idat = [&](int i) {
                const int o = i - k0;
                return o < T ? lds_ld(&sm.idw[o])
                     : (o < WIN ? 1.0 / (lds_ld(&sm.xw[o + 1]) - lds_ld(&sm.xw[o]))
                                : 1.0 / (glb_ld(&a.x[i + 1]) - glb_ld(&a.x[i])));
            };
            double Bc[U][R];
            strip_col_factors<KT, U, NW>(a, sm, k0, lane, Bc);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + lane + kWave * u;
                if (k >= n) continue;
                int cur = k;
                double xc = xk[u], xc1 = k < n - 1 ? sm.xw[lane + kWave * u + 1] : xk[u];
                double idc = sm.idw[lane + kWave * u];
                double accc = 0.0, dk = 0.0;
                const int jend = min(nrows, k - rw0 + 1);   // rows j <= k
                for (int jj = 0; jj < jend; ++jj) {
                    const double* rp = myrows + jj * RS;
                    double pr = rp[0] * Bc[u][0];
#pragma unroll
                    for (int r = 1; r < R; ++r) pr = fma(rp[r], Bc[u][r], pr);
                    if (rw0 + jj == k) pr *= 0.5;
                    dk += pr;
                    const double v = rp[R] + xk[u];
                    if (v >= xlast) {
                        // upper clip (src/aggregation_kernels.cuh:191-192)
                        accc += pr;
                    } else {
                        if (v >= xc1) {
                            int cn = cur + 1;
                            if (xat(cn + 1) <= v) {
                                const int hi = (a.inv_log_r != 0.0)
                                    ? bracket_geometric(a.x, n, v, a.log_x0, a.inv_log_r)
                                    : bracket_search(a.x, cn, n - 1, v);
                                cn = hi - 1;
                            }
                            cur = cn;
                            xc = xat(cn);
                            xc1 = xat(cn + 1);
                            idc = idat(cn);
                        }
                        const double wp = ((v - xc) * idc) * pr;
                        sink.put(cur, pr - wp);
                        sink.put(cur + 1, wp);
                    }
                }
                if (accc != 0.0) sink.put(n - 1, accc);
                if (dk != 0.0) sink.put(k, -dk);   // column k's deaths
            }
        }
        // the next piece rewrites this wave's rows
        wave_lds_sync();
    }

    // ---- far pairs of this wave's chunks: a far pair (j, k) adds
    // rate (1 - w) to bin k, rate w to bin k+1 and takes rate from bin k, so
    // bin k nets -rate w and bin k+1 +rate w ----
    double Bc[U][R];
    strip_col_factors<KT, U, NW>(a, sm, k0, lane, Bc);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int c = lane + kWave * u;
        const int k = k0 + c;
        if (k < n) {
            double accw = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) accw = fma(Bc[u][r], acc[u][r], accw);
            const double hi = accw * sm.idw[c];   // 0 for k = n-1
            if (hi != 0.0) {
                lds_add(&sm.colwin[c], -hi);
                lds_add(&sm.colwin[c + 1], hi);
            }
        }
    }
    __syncthreads();

    // ---- the next block's column loads, then flush the window ----
    if (next_q >= 0) cols.load(a, next_q);
    for (int i = tid; i < WIN; i += NTH) {
        const int k = k0 + i;
        const double v = sm.colwin[i];
        sm.colwin[i] = 0.0;
        if (k < n && v != 0.0) global_add(&a.rhs[k], v);
    }
}

}  // namespace pbe
