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

// Chunk layout of one launch.  Chunk c of column block q holds the rows
// [64 c, 64 c + 64) (clipped to the rows j <= the block's last column).
// Without the band (`band` = 0: a separate launch takes each block's diagonal
// band), block q >= 2 has (q - 1) U chunks: its rows below the band.  With the
// band, block q has (q + 1) U chunks: the rows below the band, then the band's
// lower tile (rows [(q-1)T, qT)), then the diagonal tile (rows [qT, (q+1)T)).
// Chunks [sf, se) are near chunks (the band's lower tile, and on fine grids
// the m_near strip chunks just below it), chunks [se, nc) diagonal chunks;
// the rest are far chunks, and of those the first sz are zero-birth chunks
// (with `zb` set): rows so far below the block that x_j + x_k rounds to x_k
// for every column k, so the pair's birth weight is exactly 0 and only the
// row's death remains to be added (a geometric grid of ratio r has them
// from about 54 log_r 2 bins below the block's first column down; sz(q) =
// max(0, q U - m_zero) estimates them for the cost line, the kernel tests
// every row).  The shares of the launch split a cost line on which a
// zero-birth chunk costs cz, a far chunk cf, a near chunk cn, a diagonal
// chunk cd and each block's visit (column set-up and window flush) cv.
struct StripPlan {
    long long total;   // chunks of all blocks
    int nb;            // column blocks
    int U;             // chunks per block height (T / 64)
    int band;          // 1: blocks hold their diagonal band rows too
    int m_near;        // strip chunks below the band charged as near
    int mq;            // m_near / U + 1 (formed on the host)
    int zb;            // 1: the first sz(q) chunks of block q are charged cz
    int m_zero;        // chunks below the block's first column not charged cz (>= U + m_near)
    int mz;            // m_zero / U + 1 (formed on the host)
    int cz, cf, cn, cd, cv;
};

__host__ __device__ __forceinline__ int strip_nc(const StripPlan& s, int q)
{
    return s.band ? (q + 1) * s.U : (q >= 2 ? (q - 1) * s.U : 0);
}

// first diagonal chunk of block q (nc without the band)
__host__ __device__ __forceinline__ int strip_se(const StripPlan& s, int q)
{
    return s.band ? q * s.U : strip_nc(s, q);
}

// first near chunk of block q
__host__ __device__ __forceinline__ int strip_sf(const StripPlan& s, int q)
{
    const int st = q >= 1 ? (q - 1) * s.U : 0;
    return st > s.m_near ? st - s.m_near : 0;
}

// first far chunk of block q charged cf (the chunks before it are charged cz)
__host__ __device__ __forceinline__ int strip_sz(const StripPlan& s, int q)
{
    const int st = q * s.U - s.m_zero;
    return s.zb && st > 0 ? st : 0;
}

__host__ __device__ __forceinline__ long long tri(long long m) { return m > 0 ? m * (m + 1) / 2 : 0; }

// Sums over the blocks before q of nc, se and sf, and the number of visits.
__host__ __device__ __forceinline__ long long strip_prefix(const StripPlan& s, int q)
{
    return s.band ? (long long)s.U * q * (q + 1) / 2 : s.U * tri(q - 2);
}

__host__ __device__ __forceinline__ long long strip_se_sum(const StripPlan& s, int q)
{
    return s.band ? (long long)s.U * q * (q - 1) / 2 : strip_prefix(s, q);
}

__host__ __device__ __forceinline__ long long strip_sf_sum(const StripPlan& s, int q)
{
    // sum over a = q' - 1 in [0, q - 2] of max(0, U a - m_near)
    const long long A = q - 2, a0 = s.mq;
    if (A < a0) return 0;
    return (long long)s.U * (A * (A + 1) / 2 - a0 * (a0 - 1) / 2) -
           (long long)s.m_near * (A - a0 + 1);
}

__host__ __device__ __forceinline__ long long strip_sz_sum(const StripPlan& s, int q)
{
    // sum over q' in [0, q - 1] of max(0, U q' - m_zero)
    if (!s.zb) return 0;
    const long long A = q - 1, a0 = s.mz;
    if (A < a0) return 0;
    return (long long)s.U * (A * (A + 1) / 2 - a0 * (a0 - 1) / 2) -
           (long long)s.m_zero * (A - a0 + 1);
}

__host__ __device__ __forceinline__ long long strip_cost_start(const StripPlan& s, int q)
{
    const long long nc = strip_prefix(s, q), se = strip_se_sum(s, q), sf = strip_sf_sum(s, q);
    const long long sz = strip_sz_sum(s, q);
    const long long visits = s.band ? q : (q > 2 ? q - 2 : 0);
    return s.cv * visits + s.cz * sz + s.cf * (sf - sz) + s.cn * (se - sf) + s.cd * (nc - se);
}

__host__ __device__ __forceinline__ long long strip_total(const StripPlan& s)
{
    return strip_prefix(s, s.nb);
}

__host__ __device__ __forceinline__ long long strip_cost_total(const StripPlan& s)
{
    return strip_cost_start(s, s.nb);
}

// ceil(a / b) for 0 <= a < 2^53 and 0 < b <= 2^20, through a double
// quotient: a / b lies at least 1/b from any integer it does not equal, far
// more than the quotient's rounding, so the result is exact (a 64-bit
// integer division is a long software sequence on the GPU).
__host__ __device__ __forceinline__ long long ceil_div(long long a, long long b)
{
    return (long long)ceil((double)a / (double)b);
}

// Cost of block q's visit and chunks.
__host__ __device__ __forceinline__ long long strip_block_cost(const StripPlan& s, int q)
{
    const int nc = strip_nc(s, q), se = strip_se(s, q), sf = strip_sf(s, q), sz = strip_sz(s, q);
    return (nc > 0 ? s.cv : 0) + (long long)s.cz * sz + (long long)s.cf * (sf - sz) +
           (long long)s.cn * (se - sf) + (long long)s.cd * (nc - se);
}

// Block whose cost range holds f (largest q < nb with cost_start(q) <= f).
// Past the first few blocks a block costs A + B q with B = cf U (cz U once
// zero-birth chunks are charged), so the
// start of block q is close to A q + B q (q - 1) / 2: a square root gives
// the block to within a step or two, then exact comparisons settle it.
__host__ __device__ __forceinline__ int strip_block_at_cost(const StripPlan& s, long long f)
{
    const double B = (double)(s.zb ? s.cz : s.cf) * s.U;
    const int qr = s.nb - 1;
    const double A = (double)strip_block_cost(s, qr) - B * qr;
    const double b = A - 0.5 * B;
    double g = (sqrt(b * b + 2.0 * B * (double)f) - b) / B;
    int q = g > 0.0 ? (int)g : 0;
    if (q > qr) q = qr;
    int steps = 0;
    while (q > 0 && strip_cost_start(s, q) > f && ++steps < 8) --q;
    while (q < qr && strip_cost_start(s, q + 1) <= f && ++steps < 16) ++q;
    if (steps >= 8) {
        // far from the estimate (small problems): bisect
        int lo = 0, hi = qr;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (strip_cost_start(s, mid) <= f) lo = mid; else hi = mid - 1;
        }
        q = lo;
    }
    return q;
}

// First chunk (q, c) whose start on the cost line is at or after f;
// (nb, 0) past the end.  Chunk c of block q starts at
// strip_cost_start(q) + cv + the cost of chunks [0, c).
__host__ __device__ __forceinline__ void strip_chunk_at_cost(const StripPlan& s, long long f,
                                                            int& q, int& c)
{
    if (f >= strip_cost_total(s)) {
        q = s.nb;
        c = 0;
        return;
    }
    q = strip_block_at_cost(s, f);
    const int nc = strip_nc(s, q), se = strip_se(s, q), sf = strip_sf(s, q), sz = strip_sz(s, q);
    const long long off = f - strip_cost_start(s, q) - (nc > 0 ? s.cv : 0);
    const long long e0 = (long long)s.cz * sz, e1 = e0 + (long long)s.cf * (sf - sz),
                    e2 = e1 + (long long)s.cn * (se - sf);
    long long cc;
    if (off <= 0) {
        cc = 0;
    } else if (off <= e0) {
        cc = ceil_div(off, s.cz);
    } else if (off <= e1) {
        cc = sz + ceil_div(off - e0, s.cf);
    } else if (off <= e2) {
        cc = sf + ceil_div(off - e1, s.cn);
    } else {
        cc = se + ceil_div(off - e2, s.cd);
    }
    if (cc >= nc) {
        q += 1;
        cc = 0;
    }
    c = (int)cc;
}

// Rows of a visit of block q split by cost: row j weighs cz below row 64 sz,
// cf below row 64 sf, cn below row 64 se and cd above.  row_at(t) is the
// first row whose cost offset from r0 is at least t; cost(j) the offset of
// row j.
struct VisitRows {
    int r0, r1, b0, b1, b2, wz, wf, wn, wd, c0, c1, c2;
    __host__ __device__ static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
    __host__ __device__ static int ceil_i(int a, int b) { return (int)ceil((double)a / (double)b); }
    __host__ __device__ __forceinline__ VisitRows(const StripPlan& s, int q, int r0_, int r1_)
    {
        r0 = r0_;
        r1 = r1_;
        b0 = clampi(kStripRows * strip_sz(s, q), r0, r1);
        b1 = clampi(kStripRows * strip_sf(s, q), r0, r1);
        b2 = clampi(kStripRows * strip_se(s, q), r0, r1);
        wz = s.cz;
        wf = s.cf;
        wn = s.cn;
        wd = s.cd;
        c0 = wz * (b0 - r0);
        c1 = c0 + wf * (b1 - b0);
        c2 = c1 + wn * (b2 - b1);
    }
    __host__ __device__ __forceinline__ int total() const { return c2 + wd * (r1 - b2); }
    __host__ __device__ __forceinline__ int cost(int j) const
    {
        return j <= b0 ? wz * (j - r0)
                       : (j <= b1 ? c0 + wf * (j - b0)
                                  : (j <= b2 ? c1 + wn * (j - b1) : c2 + wd * (j - b2)));
    }
    __host__ __device__ __forceinline__ int row_at(int t) const
    {
        int j;
        if (t <= c0 && c0 > 0) j = r0 + ceil_i(t, wz);
        else if (t <= c1) j = b0 + ceil_i(t - c0, wf);
        else if (t <= c2) j = b1 + ceil_i(t - c1, wn);
        else j = b2 + ceil_i(t - c2, wd);
        return clampi(j, r0, r1);
    }
};

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
    double idw[WIN];                     // 1/(x_{i+1} - x_i) of the window bins
    double colc[T];                      // cube roots of the block's pivots
    double colic[T];                     // and their reciprocals
    double coln[T];                      // populations of the block's columns
    double sk[Kern<KT>::R][T + 1];       // suffix sums of B_r over the block's columns
    int prog[NWV];                       // per wave: cost of the visit's rows done so far
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
    double xv[NW], xn[NW], nv[NT];   // as loaded: apply valid_* before use

    // Unconditional loads, indices clamped into the arrays: a load under a
    // branch makes the compiler wait for it (and for every vector-memory
    // operation before it, the window's flush atomics included) at once.
    __device__ __forceinline__ void load(const AggArgs& a, int q)
    {
        const int tid = threadIdx.x, n = a.n, k0 = q * T;
#pragma unroll
        for (int t = 0; t < NW; ++t) {
            const int b = k0 + tid + NTH * t;
            xv[t] = a.x[min(b, n - 1)];
            xn[t] = a.x[min(b + 1, n - 1)];
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) nv[t] = a.N[min(k0 + tid + NTH * t, n - 1)];
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
    // the pivot after window pivot t (inf past the window or the grid)
    __device__ __forceinline__ double next_pivot(int n, int q, int t) const
    {
        const int i = threadIdx.x + NTH * t, b = q * T + i;
        return (i < WIN && b + 1 < n) ? xn[t] : __builtin_inf();
    }
};

// Inclusive prefix sum over the 64 lanes of a wave (lane i gets the sum of
// lanes 0..i): shifts by 1, 2, 4, 8 within rows of 16 lanes, then each row's
// total into the rows above it.
__device__ __forceinline__ double wave_prefix(double v)
{
    v += dpp_move<0x111, 0xF>(v);   // row_shr:1
    v += dpp_move<0x112, 0xF>(v);   // row_shr:2
    v += dpp_move<0x114, 0xF>(v);   // row_shr:4
    v += dpp_move<0x118, 0xF>(v);   // row_shr:8
    v += dpp_move<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ double lane_value(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sk[r][c] = sum over the block's columns c' >= c of B_r(c') (sk[r][T] = 0),
// by one wave: lane l of group g takes column T - 1 - (l + 64 g), so a prefix
// over the lanes is a suffix over the columns, and the groups before g hold
// the columns above.
template <int KT, int U, int NW>
__device__ __forceinline__ void strip_suffix_sums(const AggArgs& a, StripSmem<KT, U, NW>& sm,
                                                  const int k0, const int lane)
{
    constexpr int R = Kern<KT>::R;
    constexpr int T = StripSmem<KT, U, NW>::T;
    double carry[R];
#pragma unroll
    for (int r = 0; r < R; ++r) carry[r] = 0.0;
#pragma unroll
    for (int g = 0; g < U; ++g) {
        const int c = T - 1 - (lane + kWave * g);
        double f[R];
        Kern<KT>::col(f, sm.xw[c], sm.colc[c], sm.colic[c]);
        const double Nk = k0 + c < a.n ? sm.coln[c] : 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double p = wave_prefix(k0 + c < a.n ? Nk * f[r] : 0.0);
            sm.sk[r][c] = p + carry[r];
            carry[r] += lane_value(p, kWave - 1);
        }
    }
    if (lane < R) sm.sk[lane][T] = 0.0;
}

// Half the spacing of the doubles at x, 2^(e - 53) for a normal positive x
// of exponent e (x in [2^e, 2^(e+1))); 0 for anything else (zero, negative,
// subnormal, too small for the half spacing to be a normal double, inf,
// NaN).  Any 0 <= y < half_ulp(x) leaves x + y == x in round-to-nearest, and
// so for every larger pivot too.
__host__ __device__ __forceinline__ double half_ulp(double x)
{
    long long b;
    __builtin_memcpy(&b, &x, sizeof b);
    const int eb = (int)((b >> 52) & 0x7ff);
    if (b <= 0 || eb >= 0x7ff || eb <= 53) return 0.0;
    const long long hb = (long long)(eb - 53) << 52;
    double h;
    __builtin_memcpy(&h, &hb, sizeof h);
    return h;
}

// One visit of column block q by the workgroup: chunks [c_lo, c_hi).  On
// entry `cols` holds block q's column data and the LDS window is zero; on exit
// `cols` holds block next_q's (if next_q >= 0, loaded before the window's
// global atomics so those loads do not wait behind them) and this thread's
// window bins are flushed and zero again.
//
// Rows are split over the waves by cost (a near row costs more than a far
// one).  Each wave takes its rows in pieces of up to 128; a piece's leading
// rows that are far from every column of the block run the far-pair loop,
// the rest the near-pair path.  Deaths: a row below the block loses
// sum_r A_r(j) TK_r (TK_r: the block's column sums); a row of the diagonal
// tile, sum_r A_r(j) sk_r(j + 1) (the columns above it); the near path adds
// each column's losses to its pairs explicitly, the diagonal pair's full
// rate included; the far loop nets a column's losses into its births.
template <int KT, int U, int NW = StripWaves<U>::value>
__device__ __forceinline__ void agg_strip_visit(const AggArgs& a, const StripPlan& plan,
                                                const int q, const int c_lo,
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
        // 1/(x_{i+1} - x_i) of the window bins (0 from bin n-1 on)
        if (i < WIN)
            sm.idw[i] = k0 + i < n - 1 ? 1.0 / (cols.next_pivot(n, q, t) - cols.pivot(n, q, t))
                                       : 0.0;
    }
    // the block's columns: populations, cube roots
#pragma unroll
    for (int t = 0; t < Cols::NT; ++t) {
        const int i = tid + NTH * t, k = k0 + i;
        if (i < T) {
            const double xk_ = cols.pivot(n, q, t);
            sm.coln[i] = cols.pop(n, q, t);
            double cb = 0.0, icb = 0.0;
            if (k < n) pivot_roots<KT>(xk_, cb, icb);
            sm.colc[i] = cb;
            sm.colic[i] = icb;
        }
    }
    __syncthreads();
    // a visit holding diagonal-tile rows needs the column suffix sums
    const int rmax = min(n, k0 + T);   // rows j <= last column of the block
    const int vr0 = c_lo * kStripRows, vr1 = min(c_hi * kStripRows, rmax);
    if (plan.band && vr1 > k0) {
        if (wv == 0) strip_suffix_sums<KT, U, NW>(a, sm, k0, lane);
        __syncthreads();
    }

    double xk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xk[u] = sm.xw[lane + kWave * u];
    const double hz = half_ulp(sm.xw[0]);
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
    wave_lds_sync();

    // This wave's rows: an equal share of the visit's cost.  It takes them in
    // pieces of up to 128, two rows per lane, whose factor computations
    // (cube-root chains) run interleaved.  Every 32 rows a wave posts its
    // progress (cost done) in LDS and sets its issue priority from its lead
    // over the workgroup's mean (ahead: low, behind: high): the SIMD
    // otherwise favours its oldest wave, which would finish first while the
    // younger one on the same SIMD lags (measured: 20 us against 30 us) and
    // the workgroup waits for it at the end of the visit.  This lane's rows
    // of the next piece are loaded one piece ahead, before the current piece
    // issues its global atomics: a load issued after an atomic would wait for
    // it (vector memory counts retire in order).
    constexpr int PR = kStripPieceRows / kWave;   // rows per lane in a piece
    // (a visit of far rows only splits them evenly: the same split, found
    // with less arithmetic)
    const VisitRows vrows(plan, q, vr0, vr1);
    const bool far_only = vrows.b1 == vr1 && vrows.b0 == vr0;
    const int vcost = vrows.total();
    const int w_lo = far_only ? vr0 + (vr1 - vr0) * wv / Cols::NWV
                              : vrows.row_at(vcost * wv / Cols::NWV);
    const int w_hi = far_only ? vr0 + (vr1 - vr0) * (wv + 1) / Cols::NWV
                              : vrows.row_at(vcost * (wv + 1) / Cols::NWV);
    const int w_cost0 = vrows.cost(w_lo);
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
    auto pace = [&](int row) {
        const int done = far_only ? (row - w_lo) * plan.cf : vrows.cost(row) - w_cost0;
        if (lane == 0) sm.prog[wv] = done;
        wave_lds_sync();
        int tot = 0;
#pragma unroll
        for (int w = 0; w < Cols::NWV; ++w) tot += sm.prog[w];
        const int lead = done * Cols::NWV - tot;   // NWV x (own cost - mean cost)
        const int L = 16 * Cols::NWV * plan.cf;    // 16 far rows
        set_wave_priority(lead > L ? 0 : lead > 0 ? 1 : lead > -L ? 2 : 3);
    };
    double xj_next[PR], Nj_next[PR];
    load_rows(w_lo, xj_next, Nj_next);
    for (int rw0 = w_lo; rw0 < w_hi; rw0 += kStripPieceRows) {
        pace(rw0);
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
                // death of row j against the block's columns above it
                // (separable), issued after the next piece's loads, so the
                // pair loop below hides its latency
                double dth = 0.0;
                if (j < k0) {
#pragma unroll
                    for (int r = 0; r < R; ++r) dth = fma(av[t][r], TK[r], dth);
                    global_add(&a.rhs[j], -dth);
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) dth = fma(av[t][r], sm.sk[r][j - k0 + 1], dth);
                    lds_add(&sm.colwin[j - k0], -dth);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) myrows[(lane + kWave * t) * RS + r] = av[t][r];
            myrows[(lane + kWave * t) * RS + R] = xj[t];
        }
        wave_lds_sync();

        const int nrows = rw1 - rw0;
        // The piece's leading rows below the block that are far from every
        // column (x increases, so they are a prefix): nf of them.
        auto row_far = [&](int jj) {
            const double xr = myrows[jj * RS + R];
            bool ok = true;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + lane + kWave * u;
                ok = ok && (k >= n - 1 || (xr + xk[u]) < sm.xw[lane + kWave * u + 1]);
            }
            return __all(ok);
        };
        // The piece's leading zero-birth rows: 0 <= x_j < half an ulp of the
        // block's first pivot x_k0 (so of every pivot of the block), where
        // x_j + x_k rounds to x_k and the far loop would add A_r(j) * 0 to
        // every accumulator; their deaths are added above.  Counted on the
        // rows in registers (x increases, so they are a prefix), rounded down
        // to the loop's four-row step.
        int fz = 0;
        if (hz > 0.0) {
            bool neg = false;
            int cnt = 0;
#pragma unroll
            for (int t = 0; t < PR; ++t) {
                const bool valid = rw0 + lane + kWave * t < rw1;
                neg = neg || (valid && !(xj[t] >= 0.0));
                cnt += __popcll(__ballot(valid && xj[t] >= 0.0 && xj[t] < hz));
            }
            fz = __any(neg) ? 0 : (cnt & ~3);
        }
        const int nlow = min(nrows, max(0, k0 - rw0));
        int nf;
        if (nlow == 0) {
            nf = 0;
        } else if (row_far(nlow - 1)) {
            nf = nlow;
        } else {
            int lo = 0, hi = nlow - 1;   // rows [0, lo) far; row hi is not
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (row_far(mid)) lo = mid + 1; else hi = mid;
            }
            nf = lo;
        }
        fz = min(fz, nf & ~3);
        if (nf > fz) {
            // rows [0, nf): every pair lands in (k, k+1), or clips at column
            // n-1 (taken as weight w = 0 there: 1/(x_{k+1} - x_k) is 0 for
            // k = n-1).  Rows in pairs, the second row's update of an
            // accumulator U*R instructions after the first's.  The LDS reads
            // of the next pair are issued before the FMAs of this one,
            // alternating two register sets, so their latency hides behind
            // the FMAs (past the last row they read the wave's two spare
            // rows).
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
            ld2(fz, pa0, pa1);
            int jj = fz;
            for (; jj + 3 < nf; jj += 4) {
                if ((jj & 31) == 28) pace(rw0 + jj);
                ld2(jj + 2, pb0, pb1);
                __builtin_amdgcn_sched_barrier(0);
                fma2(pa0, pa1);
                ld2(jj + 4, pa0, pa1);
                __builtin_amdgcn_sched_barrier(0);
                fma2(pb0, pb1);
            }
            // 0 to 3 rows left; rows jj and jj + 1 are in pa
            if (jj + 1 < nf) {
                fma2(pa0, pa1);
                jj += 2;
            }
            if (jj < nf) {
                const double* r0 = myrows + jj * RS;
                const double x0 = r0[R];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double w0 = (x0 + xk[u]) - xk[u];
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[u][r] = fma(r0[r], w0, acc[u][r]);
                }
            }
        }
        if (nf < nrows) {
            // Near pairs (rows [nf, nrows), the diagonal tile's among them):
            // a lane keeps the exact bracket (cur, cur+1) of v = x_j + x_k
            // for its column; rows run upward, so it only moves up.  The
            // pairs of one bracket are summed in registers (their rate and
            // weighted rate) and added to the LDS window (global atomics
            // beyond it) when the bracket moves.  One column at a time keeps
            // a single bracket live.
            auto xat = [&](int i) { return (i - k0 <= WIN) ? lds_ld(&sm.xw[i - k0]) : glb_ld(&a.x[i]); };
            auto idat = [&](int i) {
                const int o = i - k0;
                return o < WIN ? lds_ld(&sm.idw[o]) : 1.0 / (glb_ld(&a.x[i + 1]) - glb_ld(&a.x[i]));
            };
            // one loop body for the U columns (not unrolled: this path's
            // code stays small next to the far loop's)
#pragma unroll 1
            for (int u = 0; u < U; ++u) {
                const int c = lane + kWave * u;
                const int k = k0 + c;
                const int jend = k < n ? min(nrows, k - rw0 + 1) : 0;   // rows j <= k
                if (!__any(jend > nf)) continue;
                // this column's factors B_r(k)
                const double xku = sm.xw[c];
                double Bu[R];
                {
                    double f[R];
                    K::col(f, xku, sm.colc[c], sm.colic[c]);
                    const double Nk = k < n ? sm.coln[c] : 0.0;
#pragma unroll
                    for (int r = 0; r < R; ++r) Bu[r] = Nk * f[r];
                }
                // bracket (cur, cur + 1) with its pivots and 1/(x_{cur+1} -
                // x_cur), and the next bracket's pivot x_{cur+2} and
                // reciprocal, so that a move by one bin waits for no load
                int cur = k;
                double xc = xku, xc1 = sm.xw[c + 1], idc = sm.idw[c];
                double xc2 = k + 2 < n ? xat(k + 2) : __builtin_inf();
                double idn = k + 1 < n - 1 ? idat(k + 1) : 0.0;
                double accr = 0.0, accw = 0.0, accc = 0.0, dk = 0.0;
                // the rows' factors, the next row's loaded before this one's use
                double rn[R + 1];
                {
                    const double* r0 = myrows + nf * RS;
#pragma unroll
                    for (int r = 0; r <= R; ++r) rn[r] = r0[r];
                }
                for (int jj = nf; jj < jend; ++jj) {
                    double rw[R + 1];
#pragma unroll
                    for (int r = 0; r <= R; ++r) rw[r] = rn[r];
                    {
                        const double* r1 = myrows + (jj + 1) * RS;   // spare row past the end
#pragma unroll
                        for (int r = 0; r <= R; ++r) rn[r] = r1[r];
                    }
                    double rf = rw[0] * Bu[0];
#pragma unroll
                    for (int r = 1; r < R; ++r) rf = fma(rw[r], Bu[r], rf);
                    const bool dg = rw0 + jj == k;
                    const double pr = dg ? 0.5 * rf : rf;
                    dk += rf;   // column k's loss: the diagonal pair's 2 x rate/2 too
                    const double v = rw[R] + xku;
                    if (v >= xlast) {
                        // upper clip (src/aggregation_kernels.cuh:191-192)
                        accc += pr;
                        continue;
                    }
                    if (v >= xc1) {
                        if (accr != 0.0 || accw != 0.0) {
                            sink.put(cur, accr - accw);
                            sink.put(cur + 1, accw);
                        }
                        if (v < xc2) {
                            cur += 1;
                            xc = xc1;
                            xc1 = xc2;
                            idc = idn;
                        } else {
                            int cn;
                            if (v < lds_ld(&sm.xw[WIN])) {
                                cn = k0 + window_bracket_lo(sm.xw, cur + 2 - k0, WIN, v);
                            } else {
                                const int hb = (a.inv_log_r != 0.0)
                                    ? bracket_geometric(a.x, n, v, a.log_x0, a.inv_log_r)
                                    : bracket_search(a.x, cur + 2, n - 1, v);
                                cn = hb - 1;
                            }
                            cur = cn;
                            xc = xat(cn);
                            xc1 = xat(cn + 1);
                            idc = idat(cn);
                        }
                        xc2 = cur + 2 < n ? xat(cur + 2) : __builtin_inf();
                        idn = cur + 1 < n - 1 ? idat(cur + 1) : 0.0;
                        accr = 0.0;
                        accw = 0.0;
                    }
                    accr += pr;
                    accw = fma((v - xc) * idc, pr, accw);
                }
                if (accr != 0.0 || accw != 0.0) {
                    sink.put(cur, accr - accw);
                    sink.put(cur + 1, accw);
                }
                if (accc != 0.0) sink.put(n - 1, accc);
                if (dk != 0.0) sink.put(k, -dk);   // column k's losses
            }
        }
        // the next piece rewrites this wave's rows
        wave_lds_sync();
    }

    // ---- far pairs of this wave's rows: a far pair (j, k) adds rate (1 - w)
    // to bin k, rate w to bin k+1 and takes rate from bin k, so bin k nets
    // -rate w and bin k+1 +rate w ----
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
