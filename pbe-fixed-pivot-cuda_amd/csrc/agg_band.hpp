// agg_band.hpp -- the diagonal band of one column block as a diagonal sweep
// (device code; the launcher is in agg_kernels.hip).
//
// The band of column block q (T = 256 columns [k0, k0 + T)) holds its pairs
// with rows j in [max(0, k0 - T), k]: every near pair of a grid whose
// near-pair reach is below T, plus the far pairs in between.  Lane = column
// k (one per thread, 256 threads).  The pairs of a column are walked by their
// diagonal offset e = k - j from large to small, so v = x_j + x_k grows and
// the lane's bracket (cur, cur + 1) only moves up -- and since neighbouring
// columns step through the same offsets together, on a geometric grid their
// brackets move at the same step: the bracket update is a uniform branch
// rather than one taken by a few lanes every row.
//
// A work item is one block and one segment of offsets; the segment is walked
// as four interleaved streams (four independent FP64 chains per lane).  Row
// factors A_r(j) and x_j of the band's rows sit in LDS row by row, so a
// lane's gather of row j = k - e is three 16-byte reads.  Per-bracket sums are added to an LDS window of target bins, then
// flushed with global atomics.  The item of segment 0 also adds the band's
// deaths (separable, as in the tile body).
#pragma once

#include "pbe_device.hpp"
#include "pbe_internal.hpp"
#include "agg_tile.hpp"

namespace pbe {

constexpr int kBandT = 256;               // columns per block (one per thread)
constexpr int kBandRows = 2 * kBandT;     // rows of the band: block below + block
constexpr int kBandWin = kBandT + 256;    // target bins held in LDS
constexpr int kBandStreams = 4;

// Offset segments of the band: segment s of a block covers offsets
// [e[s], e[s + 1]) (block 0, whose columns have fewer rows below them, uses
// e0).  The host picks them so that segments cost about the same
// (band_segments in agg_kernels.hip).
constexpr int kBandMaxSeg = 16;
struct BandSegs {
    short e[kBandMaxSeg + 1];
    short e0[kBandMaxSeg + 1];
    // On a geometric grid of ratio r, the bracket (c, c + 1) of x_{k-e} + x_k
    // sits delta(e) = floor(log_r(1 + r^-e)) bins above k whatever k is.
    // d[s][t] is delta at the first offset of stream t of segment s (d0: the
    // same for block 0), -1 on a general grid; the kernel checks it against
    // the pivots and falls back to a search where it does not hold.
    short d[kBandMaxSeg][kBandStreams];
    short d0[kBandMaxSeg][kBandStreams];
};

template <int KT>
struct BandSmem {
    static constexpr int R = Kern<KT>::R;
    // per band row: A_r(j) then x_j, an even number of doubles so a lane
    // reads its row with 16-byte loads (a 48-byte stride at R = 5 puts the
    // 16 lanes of an LDS cycle on distinct banks)
    static constexpr int RSB = (R + 2) & ~1;
    alignas(16) double ra[kBandRows][RSB];
    double colwin[kBandWin];
    double xw[kBandWin + 1];
    double idw[kBandWin];
    double red[4][2 * R];          // per-wave TK_r (block columns), TJ_r (rows below the block)
};

// One (block, offset segment) item; every thread calls it.
template <int KT>
__device__ __forceinline__ void agg_band_sweep(const AggArgs& a, const int q, const int seg,
                                               const BandSegs& segs, BandSmem<KT>& sm)
{
    using K = Kern<KT>;
    constexpr int R = K::R;
    constexpr int T = kBandT, WIN = kBandWin;
    constexpr int NS = kBandStreams;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
    const int n = a.n;
    const int k0 = q * T;
    const int jlo = k0 - T > 0 ? k0 - T : 0;          // first band row
    const int nrows = min(n, k0 + T) - jlo;            // band rows (j <= last column)
    const int k = k0 + tid;
    const bool col = k < n;

    // ---- load phase: window pivots, band rows, this thread's column ----
    constexpr int NW = (WIN + 1 + 255) / 256;
    double xv[NW], xn[NW];
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        const int i = tid + 256 * t, b = k0 + i;
        xv[t] = (i <= WIN && b < n) ? a.x[b] : __builtin_inf();
        xn[t] = (i < WIN && b + 1 < n) ? a.x[b + 1] : __builtin_inf();
    }
    // the segment's offsets, and the band rows its pairs use: rows
    // k - e for k in the block and e in [e0, e1) (every row in segment 0,
    // which also adds the band's deaths)
    const int e0 = q == 0 ? segs.e0[seg] : segs.e[seg];
    const int e1 = q == 0 ? segs.e0[seg + 1] : segs.e[seg + 1];
    const int i_lo = seg == 0 ? 0 : max(0, k0 - e1 + 1 - jlo);
    const int i_hi = seg == 0 ? nrows : min(nrows, min(k0 + T, n) - e0 - jlo);
    double xr[2], nr[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int i = i_lo + tid + 256 * t;
        xr[t] = i < i_hi ? a.x[jlo + i] : 0.0;
        nr[t] = i < i_hi ? a.N[jlo + i] : 0.0;
    }
    const double nk = col ? a.N[k] : 0.0;
    const double xlast = a.x[n - 1];

    for (int i = tid; i < WIN; i += 256) sm.colwin[i] = 0.0;
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        const int i = tid + 256 * t;
        if (i <= WIN) sm.xw[i] = xv[t];
        if (i < WIN) sm.idw[i] = k0 + i < n - 1 ? 1.0 / (xn[t] - xv[t]) : 0.0;
    }
    // row factors A_r(j) of the band rows; partial TJ over the rows below the block
    double tj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tj[r] = 0.0;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int i = i_lo + tid + 256 * t;
        if (i < i_hi) {
            double av[R];
            double cb, icb;
            pivot_roots<KT>(xr[t], cb, icb);
            K::row(av, xr[t], cb, icb, a.coef);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                av[r] *= nr[t];
                sm.ra[i][r] = av[r];
                if (jlo + i < k0) tj[r] += av[r];
            }
            sm.ra[i][R] = xr[t];
        }
    }
    // this thread's column factors B_r(k)
    double B[R];
    const double xk = xv[0];
    {
        double cb = 0.0, icb = 0.0;
        if (col) pivot_roots<KT>(xk, cb, icb);
        double f[R];
        K::col(f, xk, cb, icb);
#pragma unroll
        for (int r = 0; r < R; ++r) B[r] = col ? nk * f[r] : 0.0;
    }
    if (seg == 0) {
        // block totals TK_r = sum_k B_r(k), TJ_r = sum over rows below the block
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double s1 = wave_sum(B[r]), s2 = wave_sum(tj[r]);
            if (lane == 0) {
                sm.red[wv][r] = s1;
                sm.red[wv][R + r] = s2;
            }
        }
    }
    __syncthreads();

    const double xk1 = sm.xw[tid + 1];   // x_{k+1} (inf past the end)
    auto xat = [&](int i) { return (i - k0 <= WIN) ? lds_ld(&sm.xw[i - k0]) : glb_ld(&a.x[i]); };
    auto idat = [&](int i) {
        const int o = i - k0;
        return o < WIN ? lds_ld(&sm.idw[o]) : 1.0 / (glb_ld(&a.x[i + 1]) - glb_ld(&a.x[i]));
    };
    auto put = [&](int i, double c) {
        const int o = i - k0;
        if (o >= 0 && o < WIN) lds_add(&sm.colwin[o], c);
        else global_add(&a.rhs[i], c);
    };
    auto jump = [&](int c, double v) {
        if (v < sm.xw[WIN]) return k0 + window_bracket_lo(sm.xw, c - k0, WIN, v);
        const int hi = (a.inv_log_r != 0.0)
            ? bracket_geometric(a.x, n, v, a.log_x0, a.inv_log_r)
            : bracket_search(a.x, c, n - 1, v);
        return hi - 1;
    };

    // ---- deaths of the band (segment 0 only) ----
    if (seg == 0) {
        double TK[R], TJ[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            TK[r] = (sm.red[0][r] + sm.red[1][r]) + (sm.red[2][r] + sm.red[3][r]);
            TJ[r] = (sm.red[0][R + r] + sm.red[1][R + r]) + (sm.red[2][R + r] + sm.red[3][R + r]);
        }
        // rows: -sum_r A_r(j) TK_r; a row of the block itself also covers the
        // column side of the block's own pairs
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int i = tid + 256 * t;
            if (i < nrows) {
                double dth = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) dth = fma(sm.ra[i][r], TK[r], dth);
                put(jlo + i, -dth);
            }
        }
        // columns: -sum_r B_r(k) TJ_r for the pairs with the rows below the block
        if (col && jlo < k0) {
            double dth = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) dth = fma(B[r], TJ[r], dth);
            lds_add(&sm.colwin[tid], -dth);
        }
    }

    // ---- births: the segment's offsets as four interleaved streams ----
    const int emax = k - jlo;                       // largest offset of this column
    const int L = (e1 - e0 + NS - 1) / NS;          // steps per stream
    // Each stream's bracket at its first pair, all four searches together:
    // a stream's first pair is where its bracket is furthest from (k, k+1),
    // and in the corner of block 0 every step starts a new column.
    int cur[NS];
    double xc[NS], xc1[NS], idc[NS], accr[NS], accw[NS], accc[NS];
    {
        int lo[NS], hi[NS];
        double vs[NS];
        bool need = false;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int eu = min(e0 + (s + 1) * L - 1, e1 - 1);   // the stream's first offset
            const int es = min(eu, emax);                       // (fewer rows in block 0)
            const bool valid = col && es >= e0 + s * L;
            vs[s] = valid ? sm.ra[k - es - jlo][R] + xk : 0.0;
            lo[s] = tid;
            // brackets beyond the window are left for the first step's search
            hi[s] = (valid && vs[s] >= xk1 && vs[s] < xlast && vs[s] < sm.xw[WIN]) ? WIN : tid + 1;
            // geometric grid: try the bracket delta(e) bins above k first
            const int dl = q == 0 ? segs.d0[seg][s] : segs.d[seg][s];
            if (hi[s] - lo[s] > 1 && dl >= 0 && es == eu && tid + dl + 1 <= WIN) {
                const int g = tid + dl;
                if (sm.xw[g] <= vs[s] && vs[s] < sm.xw[g + 1]) {
                    lo[s] = g;
                    hi[s] = g + 1;
                }
            }
            need = need || hi[s] - lo[s] > 1;
        }
        for (int step = 0; step < 10 && __any(need); ++step) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (hi[s] - lo[s] > 1) {
                    const int mid = lo[s] + ((hi[s] - lo[s]) >> 1);
                    if (sm.xw[mid] <= vs[s]) lo[s] = mid; else hi[s] = mid;
                }
            }
            need = false;
#pragma unroll
            for (int s = 0; s < NS; ++s) need = need || hi[s] - lo[s] > 1;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            cur[s] = k0 + lo[s];
            xc[s] = sm.xw[lo[s]];
            xc1[s] = sm.xw[lo[s] + 1];
            idc[s] = sm.idw[lo[s]];
            accr[s] = 0.0;
            accw[s] = 0.0;
            accc[s] = 0.0;
        }
    }
    for (int it = 0; it < L; ++it) {
        double pr[NS], vv[NS];
        bool ok[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int e = e0 + (s + 1) * L - 1 - it;   // offsets run downward
            ok[s] = col && e < e1 && e <= emax;
            const int rr = ok[s] ? k - e - jlo : 0;
            using D2 = double __attribute__((ext_vector_type(2)));
            double rw[BandSmem<KT>::RSB];
#pragma unroll
            for (int m = 0; m < BandSmem<KT>::RSB / 2; ++m) {
                const D2 d2 = *reinterpret_cast<const D2*>(&sm.ra[rr][2 * m]);
                rw[2 * m] = d2.x;
                rw[2 * m + 1] = d2.y;
            }
            double p = rw[0] * B[0];
#pragma unroll
            for (int r = 1; r < R; ++r) p = fma(rw[r], B[r], p);
            pr[s] = e == 0 ? 0.5 * p : p;
            vv[s] = rw[R] + xk;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (!ok[s]) continue;
            const double v = vv[s];
            if (v >= xlast) {
                // upper clip (src/aggregation_kernels.cuh:191-192)
                accc[s] += pr[s];
                continue;
            }
            if (v >= xc1[s]) {
                if (accr[s] != 0.0 || accw[s] != 0.0) {
                    put(cur[s], accr[s] - accw[s]);
                    put(cur[s] + 1, accw[s]);
                }
                int c = cur[s] + 1;
                if (xat(c + 1) <= v) c = jump(c, v);
                cur[s] = c;
                xc[s] = xat(c);
                xc1[s] = xat(c + 1);
                idc[s] = idat(c);
                accr[s] = 0.0;
                accw[s] = 0.0;
            }
            accr[s] += pr[s];
            accw[s] = fma((v - xc[s]) * idc[s], pr[s], accw[s]);
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (accr[s] != 0.0 || accw[s] != 0.0) {
            put(cur[s], accr[s] - accw[s]);
            put(cur[s] + 1, accw[s]);
        }
        if (accc[s] != 0.0) put(n - 1, accc[s]);
    }
    __syncthreads();

    // ---- flush the window ----
    for (int i = tid; i < WIN; i += 256) {
        const int b = k0 + i;
        const double v = sm.colwin[i];
        if (b < n && v != 0.0) global_add(&a.rhs[b], v);
    }
}

}  // namespace pbe
