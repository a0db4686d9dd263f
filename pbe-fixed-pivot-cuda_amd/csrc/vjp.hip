// vjp.hip -- hand-written vector-Jacobian products of the right-hand side and
// the discrete adjoint of explicit RK4 (no counterpart in the reference).
//
// For a cotangent lam of R, L = lam . R(N; theta).  Per aggregation pair p =
// (j <= k) with weight s_p (1/2 on the diagonal) the contribution is
//   s_p beta_jk N_j N_k g_p,   g_p = eta_p(lam) - lam_j - lam_k,
// eta_p(lam) = (1-w) lam_lo + w lam_hi (or lam_clip for a clipped pair); the
// brackets and weights depend on the pivots only.  Hence
//   dL/dN_m       = sum_k beta_mk N_k g_{mk}
//   dL/d coef_X   = 1/2 sum_m N_m sum_k phi_X(m,k) N_k g_{mk}
// with beta = sum_X coef_X phi_X (X = beta0 | beta_br, beta_sh).  Breakage:
//   dL/dN_m  = S_m (sum_q bw_q eta(x_m t_q)(lam) - lam_m)   (guarded rows)
//   dL/dS0   = sum_m (S_m/S0) N_m (...),  dL/dalpha = sum_m S_m ln(x_m/v_ref) N_m (...).
//
// The kernel is an atomic-free row gather: one lane owns row m, the four
// waves of a workgroup take a quarter of the columns k each and are combined
// in a fixed order; a finish pass (one wave per row) sums the column chunks
// and the breakage gather in a fixed order, so results are bitwise
// repeatable.  Per-row parameter partials are accumulated into a per-row
// array (each element written by one lane) and reduced in a fixed order at
// the end.
#include <algorithm>
#include <mutex>

#include "batch_plan.hpp"
#include "pbe_device.hpp"
#include "pbe_internal.hpp"

namespace pbe {

constexpr int kNP = 5;   // parameter gradient slots: beta0, beta_br, beta_sh, S0, alpha

struct VjpArgs {
    const double* __restrict__ N;
    const double* __restrict__ x;
    const double* __restrict__ t_q;
    const double* __restrict__ bw_q;
    const double* __restrict__ lam;
    const double* __restrict__ tab;   // per bin: c, 1/c, 1/(x_{k+1}-x_k), 0
    double* gN;                       // out: lam . dR/dN
    double* gth_rows;                 // += per-row parameter partials [n][kNP]
    int n, nq;
    KernelCoef coef;
    double log_x0, inv_log_r;         // geometric bracket lookup (inv_log_r = 0: search)
    int sel;                          // -1: no breakage
    double S0, v_ref, alpha, v_min;
    int log_scale;                    // 1: slots 0-3 hold d/d ln(coefficient)
    // Fused RK4 adjoint stage update (lam_acc null: plain VJP):
    //   lam_acc += gN;  av_next = c1 lam_base + c2 gN (when av_next non-null)
    //   and, after the last stage of a step (lam_copy non-null), the start of
    //   the next step: lam_copy = lam_acc, av_start = c_start lam_acc
    double* lam_acc;
    double* av_next;
    const double* lam_base;
    double c1, c2;
    double* lam_copy;
    double* av_start;
    double c_start;
};

__global__ void vjp_table_kernel(const double* __restrict__ x, double* __restrict__ tab, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double c = cbrt(x[k]);
    tab[4 * k + 0] = c;
    tab[4 * k + 1] = 1.0 / c;
    tab[4 * k + 2] = (k < n - 1) ? 1.0 / (x[k + 1] - x[k]) : 0.0;
    tab[4 * k + 3] = 0.0;
}

// phi_X(m,k) of each coefficient group of kernel KT (G groups).
template <int KT> struct VjpGroups;
template <> struct VjpGroups<0> {
    static constexpr int G = 1;
    __device__ static void phi(double* f, double, double, double, double, double, double)
    { f[0] = 1.0; }
    __device__ static double coef(int, const KernelCoef& c) { return c.b0; }
    __device__ static int slot(int) { return 0; }
};
template <> struct VjpGroups<1> {
    static constexpr int G = 1;
    __device__ static void phi(double* f, double xm, double, double, double xk, double, double)
    { f[0] = xm + xk; }
    __device__ static double coef(int, const KernelCoef& c) { return c.b0; }
    __device__ static int slot(int) { return 0; }
};
template <> struct VjpGroups<2> {
    static constexpr int G = 1;
    __device__ static void phi(double* f, double xm, double, double, double xk, double, double)
    { f[0] = xm * xk; }
    __device__ static double coef(int, const KernelCoef& c) { return c.b0; }
    __device__ static int slot(int) { return 0; }
};
template <> struct VjpGroups<3> {
    static constexpr int G = 1;
    __device__ static void phi(double* f, double, double cm, double icm, double, double ck, double ick)
    { f[0] = fma(cm, ick, fma(icm, ck, 2.0)); }
    __device__ static double coef(int, const KernelCoef& c) { return c.b_br; }
    __device__ static int slot(int) { return 1; }
};
template <> struct VjpGroups<4> {
    static constexpr int G = 1;
    __device__ static void phi(double* f, double, double cm, double, double, double ck, double)
    { const double s = cm + ck; f[0] = s * s * s; }
    __device__ static double coef(int, const KernelCoef& c) { return c.b_sh; }
    __device__ static int slot(int) { return 2; }
};
template <> struct VjpGroups<5> {
    static constexpr int G = 2;
    __device__ static void phi(double* f, double, double cm, double icm, double, double ck, double ick)
    {
        f[0] = fma(cm, ick, fma(icm, ck, 2.0));
        const double s = cm + ck;
        f[1] = s * s * s;
    }
    __device__ static double coef(int g, const KernelCoef& c) { return g == 0 ? c.b_br : c.b_sh; }
    __device__ static int slot(int g) { return g == 0 ? 1 : 2; }
};

__device__ __forceinline__ double selection_rt(int sel, double v, double S0, double v_ref,
                                               double alpha, double v_min)
{
    switch (sel) {
    case 0: return selection<0>(v, S0, v_ref, alpha, v_min);
    case 1: return selection<1>(v, S0, v_ref, alpha, v_min);
    case 2: return selection<2>(v, S0, v_ref, alpha, v_min);
    default: return selection<3>(v, S0, v_ref, alpha, v_min);
    }
}

// S(v)/S0 for the selection functions (d S / d S0)
__device__ __forceinline__ double selection_shape(int sel, double v, double v_ref,
                                                  double alpha, double v_min)
{
    switch (sel) {
    case 0: return 1.0;
    case 1: return v / v_ref;
    case 2: return pow(v / v_ref, alpha);
    default: return (v > v_min) ? 1.0 : 0.0;
    }
}

template <int KT> struct VjpSmem {
    static constexpr int G = (KT >= 0) ? VjpGroups<(KT >= 0 ? KT : 0)>::G : 1;
    double part[3][kWave][G];
};

template <int KT>
constexpr int vjp_groups() { return VjpSmem<KT>::G; }

// Aggregation part of one work item: rows [m0, m0 + 64) against the columns
// [kb, ke) of the instance described by `a`, each wave a quarter of them.
// A lane keeps the exact bracket (cur, cur+1) of v = x_m + x_k, which only
// moves up as k increases, so each pair costs a compare and, when the bracket
// moves, a few loads (the same address for all lanes where every lane's pair
// lands in (k, k+1)).  The four quarters are combined in a fixed order and
// written to part[slot][m] (no atomics).
template <int KT>
__device__ __forceinline__ void vjp_part_body(const VjpArgs& a, const int m0, const int kb,
                                              const int ke, double* __restrict__ out,
                                              VjpSmem<KT>& sm)
{
    constexpr int G = VjpSmem<KT>::G;
    using VG = VjpGroups<(KT >= 0 ? KT : 0)>;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n = a.n;
    const int m = m0 + lane;
    const bool row = m < n;
    double acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = 0.0;

    const int q = (ke - kb + 3) / 4;
    const int wb = kb + wv * q, we = min(ke, wb + q);
    if (KT >= 0 && row && wb < we) {
        const double xm = a.x[m], lm = a.lam[m];
        const double cm = a.tab[4 * m + 0], icm = a.tab[4 * m + 1];
        const double xlast = a.x[n - 1], llast = a.lam[n - 1];
        // bracket of the first pair of this quarter: for a far pair one
        // compare; else on a geometric grid the log-ratio guess with its exact
        // fix-up (a few loads), or a binary search
        int cur = -1;
        double xc = 0.0, xc1 = 0.0, lc = 0.0, lc1 = 0.0, idc = 0.0;
        {
            // v0 >= x_lo with lo = max(m, wb): a far pair's bracket is
            // (lo, lo + 1), one compare
            const double v0 = xm + a.x[wb];
            const int lo0 = max(m, wb);
            if (v0 < xlast) {
                cur = v0 < a.x[lo0 + 1]
                    ? lo0
                    : (a.inv_log_r != 0.0
                           ? bracket_geometric(a.x, n, v0, a.log_x0, a.inv_log_r)
                           : bracket_search(a.x, lo0, n - 1, v0)) - 1;
                xc = a.x[cur];
                xc1 = a.x[cur + 1];
                lc = a.lam[cur];
                lc1 = a.lam[cur + 1];
                idc = a.tab[4 * cur + 2];
            }
        }
        // column data of the next pair loaded before this pair's use (the
        // index clamped into the quarter, so the loads are unconditional)
        auto colload = [&](int k, double& xk, double& Nk, double& lk, double& ck,
                           double& ick) {
            const int kk = min(k, we - 1);
            xk = a.x[kk];
            Nk = a.N[kk];
            lk = a.lam[kk];
            ck = a.tab[4 * kk + 0];
            ick = a.tab[4 * kk + 1];
        };
        double xn, Nn, ln, cn, icn;
        colload(wb, xn, Nn, ln, cn, icn);
        for (int k = wb; k < we; ++k) {
            const double xk = xn, Nk = Nn, lk = ln, ck = cn, ick = icn;
            colload(k + 1, xn, Nn, ln, cn, icn);
            const double v = xm + xk;
            double eta;
            if (v >= xlast) {
                eta = llast;                     // clipped to the last pivot
            } else {
                while (v >= xc1) {
                    ++cur;
                    xc = xc1;
                    lc = lc1;
                    xc1 = a.x[cur + 1];
                    lc1 = a.lam[cur + 1];
                    idc = a.tab[4 * cur + 2];
                }
                const double w = (v - xc) * idc;
                eta = fma(w, lc1 - lc, lc);
            }
            const double t = Nk * ((eta - lm) - lk);
            double f[G];
            VG::phi(f, xm, cm, icm, xk, ck, ick);
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g] = fma(f[g], t, acc[g]);
        }
    }
    if (wv > 0) {
#pragma unroll
        for (int g = 0; g < G; ++g) sm.part[wv - 1][lane][g] = acc[g];
    }
    __syncthreads();
    if (wv == 0 && row) {
#pragma unroll
        for (int g = 0; g < G; ++g)
            out[(size_t)G * m + g] =
                ((acc[g] + sm.part[0][lane][g]) + sm.part[1][lane][g]) + sm.part[2][lane][g];
    }
}

// Columns per chunk of the single-instance part launch (its LDS staging
// holds them) and bracket bins staged beyond the chunk.
constexpr int kVjpMaxCols = 256;
constexpr int kVjpWinExtra = 64;
constexpr int kVjpWin = kVjpMaxCols + kVjpWinExtra + 1;

struct VjpStage {
    double wx[kVjpWin], wl[kVjpWin], wid[kVjpWin];   // bins kb..: x, lam, 1/(x_{i+1}-x_i)
    double cn[kVjpMaxCols], cc[kVjpMaxCols], cic[kVjpMaxCols];   // columns: N, c, 1/c
};

// vjp_part_body with the chunk's column data and the bins just above it
// staged in LDS by the whole workgroup first (ke - kb <= kVjpMaxCols): the
// pair loop then reads a column as LDS broadcasts, and a bracket move within
// the staged bins reads LDS too (beyond them, global memory).
template <int KT>
__device__ __forceinline__ void vjp_part_body_staged(const VjpArgs& a, const int m0,
                                                     const int kb, const int ke,
                                                     double* __restrict__ out,
                                                     VjpSmem<KT>& sm, VjpStage& st)
{
    constexpr int G = VjpSmem<KT>::G;
    using VG = VjpGroups<(KT >= 0 ? KT : 0)>;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n = a.n;
    const int m = m0 + lane;
    const bool row = m < n;
    const int nw = min(ke - kb + kVjpWinExtra + 1, n - kb);   // staged bins
    for (int i = tid; i < nw; i += blockDim.x) {
        const int b = kb + i;
        st.wx[i] = a.x[b];
        st.wl[i] = a.lam[b];
        st.wid[i] = a.tab[4 * b + 2];
    }
    for (int i = tid; i < ke - kb; i += blockDim.x) {
        const int k = kb + i;
        st.cn[i] = a.N[k];
        st.cc[i] = a.tab[4 * k + 0];
        st.cic[i] = a.tab[4 * k + 1];
    }
    double acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = 0.0;
    const int q = (ke - kb + 3) / 4;
    const int wb = kb + wv * q, we = min(ke, wb + q);
    double xm = 0.0, lm = 0.0, cm = 0.0, icm = 0.0, xlast = 0.0, llast = 0.0;
    if (row) {
        xm = a.x[m];
        lm = a.lam[m];
        cm = a.tab[4 * m + 0];
        icm = a.tab[4 * m + 1];
        xlast = a.x[n - 1];
        llast = a.lam[n - 1];
    }
    __syncthreads();
    if (KT >= 0 && row && wb < we) {
        auto xat = [&](int i) { return i - kb < nw ? lds_ld(&st.wx[i - kb]) : glb_ld(&a.x[i]); };
        auto lat = [&](int i) { return i - kb < nw ? lds_ld(&st.wl[i - kb]) : glb_ld(&a.lam[i]); };
        auto iat = [&](int i) {
            return i - kb < nw ? lds_ld(&st.wid[i - kb]) : glb_ld(&a.tab[4 * i + 2]);
        };
        // bracket of the first pair of this quarter: v0 >= x_lo with
        // lo = max(m, wb), and for a far pair (the row far below the
        // column, or the column far below the row) the bracket is (lo,
        // lo + 1), one compare; else a search
        int cur = -1;
        double xc = 0.0, xc1 = 0.0, lc = 0.0, lc1 = 0.0, idc = 0.0;
        {
            const double v0 = xm + st.wx[wb - kb];
            const int lo0 = max(m, wb);
            if (v0 < xlast && v0 < xat(lo0 + 1)) {
                cur = lo0;
                xc = xat(cur);
                xc1 = xat(cur + 1);
                lc = lat(cur);
                lc1 = lat(cur + 1);
                idc = iat(cur);
            } else if (v0 < xlast) {
                int hi;
                if (v0 < st.wx[nw - 1]) {
                    int lo = wb - kb, h = nw - 1;   // wx[lo] <= v0 < wx[h]
                    while (h - lo > 1) {
                        const int mid = (lo + h) >> 1;
                        if (st.wx[mid] <= v0) lo = mid; else h = mid;
                    }
                    hi = kb + h;
                } else {
                    hi = a.inv_log_r != 0.0 ? bracket_geometric(a.x, n, v0, a.log_x0, a.inv_log_r)
                                            : bracket_search(a.x, kb + nw - 1, n - 1, v0);
                }
                cur = hi - 1;
                xc = xat(cur);
                xc1 = xat(cur + 1);
                lc = lat(cur);
                lc1 = lat(cur + 1);
                idc = iat(cur);
            }
        }
        for (int k = wb; k < we; ++k) {
            const int o = k - kb;
            const double xk = st.wx[o], Nk = st.cn[o], lk = st.wl[o];
            const double ck = st.cc[o], ick = st.cic[o];
            const double v = xm + xk;
            double eta;
            if (v >= xlast) {
                eta = llast;                     // clipped to the last pivot
            } else {
                while (v >= xc1) {
                    ++cur;
                    xc = xc1;
                    lc = lc1;
                    xc1 = xat(cur + 1);
                    lc1 = lat(cur + 1);
                    idc = iat(cur);
                }
                const double w = (v - xc) * idc;
                eta = fma(w, lc1 - lc, lc);
            }
            const double t = Nk * ((eta - lm) - lk);
            double f[G];
            VG::phi(f, xm, cm, icm, xk, ck, ick);
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g] = fma(f[g], t, acc[g]);
        }
    }
    if (wv > 0) {
#pragma unroll
        for (int g = 0; g < G; ++g) sm.part[wv - 1][lane][g] = acc[g];
    }
    __syncthreads();
    if (wv == 0 && row) {
#pragma unroll
        for (int g = 0; g < G; ++g)
            out[(size_t)G * m + g] =
                ((acc[g] + sm.part[0][lane][g]) + sm.part[1][lane][g]) + sm.part[2][lane][g];
    }
}

// gN_m and the fused adjoint stage update of row m.
__device__ __forceinline__ void vjp_store_row(const VjpArgs& a, const int m, const double gn)
{
    a.gN[m] = gn;
    if (a.lam_acc) {
        const double l = a.lam_acc[m] + gn;
        a.lam_acc[m] = l;
        if (a.av_next) a.av_next[m] = fma(a.c2, gn, a.c1 * a.lam_base[m]);
        if (a.lam_copy) {
            a.lam_copy[m] = l;
            a.av_start[m] = a.c_start * l;
        }
    }
}

// Per row: sum the column-chunk partials in order, form lam . dR/dN_m
// (aggregation + breakage) and the row's parameter partials.
template <int KT>
__device__ __forceinline__ void vjp_finish_row(const VjpArgs& a, const int m,
                                               const double* __restrict__ part, int nchunk,
                                               size_t chunk_stride)
{
    constexpr int G = VjpSmem<KT>::G;
    const int n = a.n;
    const double Nm = a.N[m], lm = a.lam[m], xm = a.x[m];
    double gn = 0.0;
    double* gr = a.gth_rows + (size_t)kNP * m;
    if constexpr (KT >= 0) {
        using VG = VjpGroups<KT>;
        double acc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] = 0.0;
        for (int c = 0; c < nchunk; ++c) {
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g] += part[c * chunk_stride + (size_t)G * m + g];
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double cf = VG::coef(g, a.coef);
            gn = fma(cf, acc[g], gn);
            gr[VG::slot(g)] += (a.log_scale ? cf : 1.0) * (0.5 * Nm * acc[g]);
        }
    }
    if (a.sel >= 0) {
        const double S = selection_rt(a.sel, xm, a.S0, a.v_ref, a.alpha, a.v_min);
        if (S != 0.0 && Nm > 2.2e-308) {
            double e = 0.0;
            for (int q = 0; q < a.nq; ++q) {
                const double v = xm * a.t_q[q];
                double et;
                if (v <= a.x[0]) et = a.lam[0];
                else if (v >= a.x[n - 1]) et = a.lam[n - 1];
                else {
                    const int hi = bracket_search(a.x, 0, n - 1, v), lo = hi - 1;
                    const double xl = a.x[lo], xh = a.x[hi];
                    const double w = (v - xl) / (xh - xl);
                    et = (1.0 - w) * a.lam[lo] + w * a.lam[hi];
                }
                e = fma(a.bw_q[q], et, e);
            }
            const double h = e - lm;
            gn = fma(S, h, gn);
            gr[3] += (a.log_scale ? a.S0 : 1.0) *
                     (selection_shape(a.sel, xm, a.v_ref, a.alpha, a.v_min) * Nm * h);
            if (a.sel == 2) gr[4] += S * log(xm / a.v_ref) * Nm * h;
        }
    }
    vjp_store_row(a, m, gn);
}

// Breakage gather of row m: sum_q bw_q eta(x_m t_q)(lam), one wave per row,
// lane q taking the quadrature points q, q + 64, ... and the wave summing
// them in the fixed butterfly order of wave_sum, so the bracket searches of a
// row run side by side.  It depends on lam and the pivots only; the finish
// pass applies S_m and the row guards.
__device__ __forceinline__ double vjp_brk_gather_wave(const VjpArgs& a, const int m)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int n = a.n;
    const double xm = a.x[m];
    // fragments below x_{m+1} are bracketed within [0, m + 1]
    const int mh = min(m + 1, n - 1);
    const double xmh = a.x[mh];
    double e = 0.0;
    for (int q = lane; q < a.nq; q += kWave) {
        const double v = xm * a.t_q[q];
        double et;
        if (v <= a.x[0]) et = a.lam[0];
        else if (v >= a.x[n - 1]) et = a.lam[n - 1];
        else {
            const int hi = a.inv_log_r != 0.0
                ? bracket_geometric(a.x, n, v, a.log_x0, a.inv_log_r)
                : bracket_search(a.x, 0, v < xmh ? mh : n - 1, v);
            const int lo = hi - 1;
            const double xl = a.x[lo], xh = a.x[hi];
            const double w = (v - xl) / (xh - xl);
            et = (1.0 - w) * a.lam[lo] + w * a.lam[hi];
        }
        e = fma(a.bw_q[q], et, e);
    }
    return wave_sum(e);
}

// Single instance: one wave per row.  Lane c adds the column-chunk partials
// c, c + 64, ..., summed over the wave in wave_sum's fixed order; the
// breakage terms of the row come precomputed in brk[4m..4m+3] (the gather
// sum_q bw_q eta(x_m t_q), S_m, dS_m/dS0 in the gradient's scale, and
// S_m ln(x_m / v_ref)).  Every load of the row, the read-modify-writes
// included, is issued before the sums that need them.
template <int KT>
__device__ __forceinline__ void vjp_finish_row_wave(const VjpArgs& a, const int m,
                                                    const double* __restrict__ part,
                                                    int nchunk, size_t chunk_stride,
                                                    const double* __restrict__ brk)
{
    constexpr int G = VjpSmem<KT>::G;
    const int lane = threadIdx.x & (kWave - 1);
    const double Nm = a.N[m], lm = a.lam[m];
    double* grow = a.gth_rows + (size_t)kNP * m;
    double gr[kNP], lacc = 0.0, lbase = 0.0;
#pragma unroll
    for (int p = 0; p < kNP; ++p) gr[p] = grow[p];
    if (a.lam_acc) lacc = a.lam_acc[m];
    if (a.av_next) lbase = a.lam_base[m];
    double b4[4] = {0.0, 0.0, 0.0, 0.0};
    if (a.sel >= 0) {
#pragma unroll
        for (int t = 0; t < 4; ++t) b4[t] = brk[4 * (size_t)m + t];
    }
    double gn = 0.0;
    if constexpr (KT >= 0) {
        using VG = VjpGroups<KT>;
        double acc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) acc[g] = 0.0;
        for (int c = lane; c < nchunk; c += kWave) {
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g] += part[c * chunk_stride + (size_t)G * m + g];
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double t = wave_sum(acc[g]);
            const double cf = VG::coef(g, a.coef);
            gn = fma(cf, t, gn);
            gr[VG::slot(g)] += (a.log_scale ? cf : 1.0) * (0.5 * Nm * t);
        }
    }
    if (a.sel >= 0) {
        const double S = b4[1];
        if (S != 0.0 && Nm > 2.2e-308) {
            const double h = b4[0] - lm;
            gn = fma(S, h, gn);
            gr[3] += b4[2] * Nm * h;
            if (a.sel == 2) gr[4] += b4[3] * Nm * h;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int p = 0; p < kNP; ++p) grow[p] = gr[p];
        a.gN[m] = gn;
        if (a.lam_acc) {
            const double l = lacc + gn;
            a.lam_acc[m] = l;
            if (a.av_next) a.av_next[m] = fma(a.c2, gn, a.c1 * lbase);
            if (a.lam_copy) {
                a.lam_copy[m] = l;
                a.av_start[m] = a.c_start * l;
            }
        }
    }
}

// Rows per breakage slice of the single-instance part launch: one per wave.
constexpr int kVjpBrkRows = 256 / kWave;

// Single instance.  Phase 1: grid (64-row blocks) x (nch column chunks, then
// nbs breakage slices of kVjpBrkRows rows each), so the breakage gathers run
// beside the aggregation chunks rather than after them.
template <int KT>
__global__ __launch_bounds__(256)
void vjp_part_kernel(VjpArgs a, double* part, int kc, int nch, double* brk_e)
{
    __shared__ VjpSmem<KT> sm;
    __shared__ VjpStage st;
    constexpr int G = VjpSmem<KT>::G;
    const int c = blockIdx.y;
    if (c >= nch) {
        // breakage terms of row m: the gather and the selection values the
        // finish pass needs (pow / log once here, beside the pair chunks)
        const int m = blockIdx.x * kWave + (c - nch) * kVjpBrkRows + (int)(threadIdx.x / kWave);
        if (m < a.n) {
            const double e = vjp_brk_gather_wave(a, m);
            if ((threadIdx.x & (kWave - 1)) == 0) {
                const double xm = a.x[m];
                const double S = selection_rt(a.sel, xm, a.S0, a.v_ref, a.alpha, a.v_min);
                double* o = brk_e + 4 * (size_t)m;
                o[0] = e;
                o[1] = S;
                o[2] = (a.log_scale ? a.S0 : 1.0) *
                       selection_shape(a.sel, xm, a.v_ref, a.alpha, a.v_min);
                o[3] = a.sel == 2 ? S * log(xm / a.v_ref) : 0.0;
            }
        }
        return;
    }
    if constexpr (KT >= 0) {
        const int kb = c * kc, ke = min(a.n, kb + kc);
        vjp_part_body_staged<KT>(a, blockIdx.x * kWave, kb, ke, part + (size_t)c * G * a.n,
                                 sm, st);
    }
}

// Phase 2: one wave per row.
template <int KT>
__global__ __launch_bounds__(256)
void vjp_finish_kernel(VjpArgs a, const double* part, int nchunk, const double* brk_e)
{
    constexpr int G = VjpSmem<KT>::G;
    const int m = blockIdx.x * (256 / kWave) + (int)(threadIdx.x / kWave);
    if (m < a.n) vjp_finish_row_wave<KT>(a, m, part, nchunk, (size_t)G * a.n, brk_e);
}

// Batched.  Items are (instance, 64-row block); every instance is cut into
// kBatchVjpChunks column chunks.  Slots 0-3 of the per-row partials hold
// d/d ln(coefficient) so that their sum over instances is the gradient of
// jointly scaled coefficients.
constexpr int kBatchVjpChunks = 8;

__device__ __forceinline__ VjpArgs instance_args(const VjpArgs& base, const BatchInstDev& d)
{
    VjpArgs a = base;
    a.N = base.N + d.off;
    a.x = base.x + d.off;
    a.lam = base.lam + d.off;
    a.tab = base.tab + 4 * (size_t)d.off;
    a.gN = base.gN + d.off;
    a.gth_rows = base.gth_rows + (size_t)kNP * d.off;
    if (base.lam_acc) a.lam_acc = base.lam_acc + d.off;
    if (base.av_next) {
        a.av_next = base.av_next + d.off;
        a.lam_base = base.lam_base + d.off;
    }
    if (base.lam_copy) {
        a.lam_copy = base.lam_copy + d.off;
        a.av_start = base.av_start + d.off;
    }
    a.n = d.n;
    a.coef = KernelCoef{d.b0, d.b_br, d.b_sh};
    a.log_x0 = d.log_x0;
    a.inv_log_r = d.inv_log_r;
    a.S0 = d.S0;
    a.v_ref = d.v_ref;
    a.alpha = d.alpha;
    a.v_min = d.v_min;
    return a;
}

template <int KT>
__global__ __launch_bounds__(256)
void vjp_batch_part_kernel(VjpArgs base, const BatchInstDev* __restrict__ inst,
                           const BatchItem* __restrict__ items, int nitems, double* part,
                           int total)
{
    __shared__ VjpSmem<KT> sm;
    constexpr int G = VjpSmem<KT>::G;
    const int nwork = nitems * kBatchVjpChunks;
    for (int it = blockIdx.x; it < nwork; it += gridDim.x) {
        const int c = it % kBatchVjpChunks;
        const BatchItem w = items[it / kBatchVjpChunks];
        const BatchInstDev d = inst[w.inst];
        const VjpArgs a = instance_args(base, d);
        const int kc = (d.n + kBatchVjpChunks - 1) / kBatchVjpChunks;
        const int kb = min(d.n, c * kc), ke = min(d.n, kb + kc);
        vjp_part_body<KT>(a, w.unit * kWave, kb, ke,
                          part + ((size_t)c * total + d.off) * G, sm);
        __syncthreads();
    }
}

template <int KT>
__global__ __launch_bounds__(256)
void vjp_batch_finish_kernel(VjpArgs base, const BatchInstDev* __restrict__ inst,
                             const BatchItem* __restrict__ items, int nitems,
                             const double* part, int total)
{
    constexpr int G = VjpSmem<KT>::G;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BatchItem w = items[it];
        const BatchInstDev d = inst[w.inst];
        const VjpArgs a = instance_args(base, d);
        const int m = w.unit * kWave + (int)threadIdx.x;
        if (threadIdx.x < kWave && m < d.n)
            vjp_finish_row<KT>(a, m, part + (size_t)G * d.off, kBatchVjpChunks,
                               (size_t)G * total);
    }
}

hipError_t enqueue_vjp_table(const double* x, double* tab, int n, hipStream_t s)
{
    hipLaunchKernelGGL(vjp_table_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, tab, n);
    return hipGetLastError();
}

// Column chunks of the single-instance VJP for n bins: about kVjpItems work
// items, at least 16 and at most kVjpMaxCols columns per chunk.
constexpr int kVjpItems = 512;
static int vjp_chunks(int n)
{
    const int nrb = (n + kWave - 1) / kWave;
    int c = (kVjpItems + nrb - 1) / nrb;
    c = std::min(c, 64);
    c = std::min(c, std::max(1, n / 16));
    c = std::max(c, (n + kVjpMaxCols - 1) / kVjpMaxCols);
    return std::max(c, 1);
}

// Single instance: nch chunks of up to 2 groups per row, then the breakage
// terms of every row (4 doubles each).
size_t vjp_part_doubles(int n, bool batched)
{
    return batched ? (size_t)kBatchVjpChunks * 2 * (size_t)n
                   : ((size_t)vjp_chunks(n) * 2 + 4) * (size_t)n;
}

template <int KT>
static void launch_vjp_single(const VjpArgs& a, double* part, hipStream_t s)
{
    const int nch = KT >= 0 ? vjp_chunks(a.n) : 0;
    const int kc = nch > 0 ? (a.n + nch - 1) / nch : 0;
    const int nbs = a.sel >= 0 ? kWave / kVjpBrkRows : 0;
    double* brk_e = part + (size_t)vjp_chunks(a.n) * 2 * (size_t)a.n;
    const dim3 b(256);
    if (nch + nbs > 0)
        hipLaunchKernelGGL(vjp_part_kernel<KT>,
                           dim3((unsigned)((a.n + kWave - 1) / kWave), (unsigned)(nch + nbs)),
                           b, 0, s, a, part, kc, nch, brk_e);
    hipLaunchKernelGGL(vjp_finish_kernel<KT>, dim3((unsigned)((a.n + 3) / 4)), b, 0, s, a,
                       (const double*)part, nch, (const double*)brk_e);
}

template <int KT>
static void launch_vjp_batch(const VjpArgs& a, const pbe_batch_plan_s* pl, double* part,
                             hipStream_t s)
{
    const dim3 b(256);
    const int nw = pl->n_row_items * kBatchVjpChunks;
    if (KT >= 0)
    hipLaunchKernelGGL(vjp_batch_part_kernel<KT>, dim3((unsigned)std::min(pl->grid, nw)), b, 0,
                       s, a, (const BatchInstDev*)pl->d_inst, (const BatchItem*)pl->d_row_items,
                       pl->n_row_items, part, pl->total);
    hipLaunchKernelGGL(vjp_batch_finish_kernel<KT>,
                       dim3((unsigned)std::min(pl->grid, pl->n_row_items)), b, 0, s, a,
                       (const BatchInstDev*)pl->d_inst, (const BatchItem*)pl->d_row_items,
                       pl->n_row_items, (const double*)part, pl->total);
}

hipError_t enqueue_vjp(const RhsProblem& p, const double* N, const double* lam,
                       const double* tab, double* gN, double* gth_rows, double* part,
                       hipStream_t s, const AdjointStageFuse* fuse)
{
    VjpArgs a{};
    if (fuse) {
        a.lam_acc = fuse->lam_acc;
        a.av_next = fuse->av_next;
        a.lam_base = fuse->lam_base;
        a.c1 = fuse->c1;
        a.c2 = fuse->c2;
        a.lam_copy = fuse->lam_copy;
        a.av_start = fuse->av_start;
        a.c_start = fuse->c_start;
    }
    a.N = N;
    a.x = p.x;
    a.t_q = p.t_q;
    a.bw_q = p.bw_q;
    a.lam = lam;
    a.tab = tab;
    a.gN = gN;
    a.gth_rows = gth_rows;
    a.n = p.n;
    if (p.agg) {
        a.log_x0 = p.agg->log_x0;
        a.inv_log_r = p.agg->inv_log_r;
    }
    int kt = -1;
    if (p.batch) {
        const pbe_batch_plan_s* pl = p.batch;
        a.nq = pl->selection >= 0 ? pl->n_quad : 0;
        a.sel = pl->selection;
        a.log_scale = 1;
        kt = pl->kernel_type;
        switch (kt) {
        case -1: launch_vjp_batch<-1>(a, pl, part, s); break;
        case 0: launch_vjp_batch<0>(a, pl, part, s); break;
        case 1: launch_vjp_batch<1>(a, pl, part, s); break;
        case 2: launch_vjp_batch<2>(a, pl, part, s); break;
        case 3: launch_vjp_batch<3>(a, pl, part, s); break;
        case 4: launch_vjp_batch<4>(a, pl, part, s); break;
        case 5: launch_vjp_batch<5>(a, pl, part, s); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    a.nq = p.brk ? p.brk->n_quad : 0;
    a.sel = p.brk ? p.brk->selection : -1;
    if (p.agg) a.coef = KernelCoef{p.agg->beta0, p.agg->beta_br, p.agg->beta_sh};
    if (p.brk) {
        a.S0 = p.brk->S0;
        a.v_ref = p.brk->v_ref;
        a.alpha = p.brk->alpha;
        a.v_min = p.brk->v_min;
    }
    kt = p.agg ? p.agg->kernel_type : -1;
    switch (kt) {
    case -1: launch_vjp_single<-1>(a, part, s); break;
    case 0: launch_vjp_single<0>(a, part, s); break;
    case 1: launch_vjp_single<1>(a, part, s); break;
    case 2: launch_vjp_single<2>(a, part, s); break;
    case 3: launch_vjp_single<3>(a, part, s); break;
    case 4: launch_vjp_single<4>(a, part, s); break;
    case 5: launch_vjp_single<5>(a, part, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Discrete adjoint of RK4.  With stage states Y1..Y4 and a_s = dL/dk_s:
//   a4 = dt/6 lam          mu4 = J4^T a4
//   a3 = dt/3 lam + dt mu4 mu3 = J3^T a3
//   a2 = dt/3 lam + dt/2 mu3, a1 = dt/6 lam + dt/2 mu2
//   lam <- lam + mu1 + mu2 + mu3 + mu4;  dL/dtheta += sum_s a_s . dR/dtheta(Y_s)
// ---------------------------------------------------------------------------
__global__ void adj_start_kernel(const double* __restrict__ lam, double* __restrict__ lam_new,
                                 double* __restrict__ av, int n, double c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double l = lam[i];
    lam_new[i] = l;
    av[i] = c * l;
}

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// gtheta[p] = sum_m rows[m][p], fixed order (one workgroup, tree per slot)
__global__ __launch_bounds__(256) void gth_reduce_kernel(const double* __restrict__ rows,
                                                         double* __restrict__ out, int n)
{
    __shared__ double red[256];
    for (int p = 0; p < kNP; ++p) {
        double s = 0.0;
        for (int m = threadIdx.x; m < n; m += 256) s += rows[(size_t)kNP * m + p];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[p] = red[0];
        __syncthreads();
    }
}

static inline dim3 grid1(int n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t enqueue_gtheta_reduce(const double* rows, double* out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(gth_reduce_kernel, dim3(1), dim3(256), 0, s, rows, out, n);
    return hipGetLastError();
}

hipError_t enqueue_copy(const double* src, double* dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(copy_kernel, grid1(n), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// every > 0: a checkpoint every `every` steps and the stage states of one
// segment, recomputed in the reverse pass.  every == 0: the forward pass keeps
// the four stage states of every step (4 steps n doubles) and the reverse
// pass recomputes nothing.
AdjointLayout adjoint_layout(int n, int steps, int every)
{
    AdjointLayout L{};
    const size_t nn = (size_t)n;
    L.store_all = every == 0;
    L.nseg = L.store_all ? 1 : (steps + every - 1) / every;
    const size_t seg_steps = L.store_all ? (size_t)steps : (size_t)every;
    size_t off = 0;
    L.ckpt = off;  off += L.store_all ? 0 : nn * (size_t)L.nseg;
    L.seg = off;   off += nn * 4 * seg_steps;
    L.seg2 = off;  off += L.store_all ? 0 : nn * 4 * seg_steps;
    L.tab = off;   off += 4 * nn;
    L.lam = off;   off += nn;
    L.lam2 = off;  off += nn;
    L.av = off;    off += 2 * nn;
    L.mu = off;    off += nn;
    L.rk = off;    off += 3 * nn;
    L.rk2 = off;   off += L.store_all ? 0 : 3 * nn;
    L.mu2 = off;   off += L.store_all ? 0 : nn;
    L.gth = off;   off += (size_t)kNP * nn;
    L.part = off;  off += std::max(vjp_part_doubles(n, false), vjp_part_doubles(n, true));
    L.total = off; // doubles
    return L;
}

hipError_t enqueue_adjoint_forward(const RhsProblem& p, double* N, double dt, int steps,
                                   int every, double* ws, hipStream_t s)
{
    const AdjointLayout L = adjoint_layout(p.n, steps, every);
    double* rk = ws + L.rk;
    const size_t nn = (size_t)p.n;
    hipError_t e = enqueue_fill_zero(rk, p.n, s);
    for (int i = 0; e == hipSuccess && i < steps; ++i) {
        if (L.store_all) {
            // the first stage kernel keeps Y1 = N in Yi
            double* Yi = ws + L.seg + nn * 4 * (size_t)i;
            double* st[3] = {Yi + nn, Yi + 2 * nn, Yi + 3 * nn};
            e = enqueue_rk4_step_io(p, N, N, dt, rk, st, Yi, s);
            continue;
        }
        if (i % every == 0) e = enqueue_copy(N, ws + L.ckpt + nn * (i / every), p.n, s);
        if (e == hipSuccess) e = enqueue_rk4_step_stages(p, N, dt, rk, nullptr, s);
    }
    return e;
}

// The library's second stream on the current device (default priority,
// created on first use and kept): the checkpointed adjoint recomputes the
// next segment there while the reverse sweep of the current one runs on the
// caller's stream.
static hipError_t side_stream(hipStream_t* out)
{
    static std::mutex mu;
    static hipStream_t streams[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!streams[dev]) {
        e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
        if (e != hipSuccess) {
            streams[dev] = nullptr;
            return e;
        }
    }
    *out = streams[dev];
    return hipSuccess;
}

hipError_t enqueue_adjoint_backward(const RhsProblem& p, const double* gN_final, double* gN0,
                                    double* gtheta, double dt, int steps, int every,
                                    double* ws, hipStream_t s)
{
    const int n = p.n;
    const size_t nn = (size_t)n;
    const AdjointLayout L = adjoint_layout(n, steps, every);
    double* tab = ws + L.tab;
    double* lam = ws + L.lam;
    double* lam2 = ws + L.lam2;
    double* av0 = ws + L.av;
    double* av1 = ws + L.av + nn;
    double* mu = ws + L.mu;
    double* gth = ws + L.gth;
    double* part = ws + L.part;
    hipError_t e;
    if ((e = enqueue_vjp_table(p.x, tab, n, s)) != hipSuccess) return e;
    if ((e = enqueue_fill_zero(gth, kNP * n, s)) != hipSuccess) return e;
    if ((e = enqueue_copy(gN_final, lam, n, s)) != hipSuccess) return e;
    const dim3 g = grid1(n), b(256);

    // Recompute segment sg from its checkpoint into Y (every stage state),
    // on stream ss with integrator scratch rkw (k zero on entry and exit)
    // and muw.
    auto recompute = [&](int sg, double* Y, double* rkw, double* muw,
                         hipStream_t ss) -> hipError_t {
        const int s0 = sg * every, s1 = min(steps, s0 + every);
        hipError_t er = enqueue_copy(ws + L.ckpt + nn * sg, Y, n, ss);
        for (int i = s0; er == hipSuccess && i < s1; ++i) {
            double* Yi = Y + nn * 4 * (size_t)(i - s0);
            double* st[3] = {Yi + nn, Yi + 2 * nn, Yi + 3 * nn};
            // the step goes from Yi to Y1 of step i+1; the segment's last
            // step only needs its stage states
            double* next = (i + 1 < s1) ? Yi + 4 * nn : muw;
            er = enqueue_rk4_step_io(p, Yi, next, dt, rkw, st, nullptr, ss);
        }
        return er;
    };
    // Reverse sweep over steps [s0, s1) with stage states Y; each VJP's
    // finish pass also adds mu_s to the new lam and forms the next stage's
    // cotangent, and the last stage's finish starts the next step (lam2 <-
    // lam, av0 <- dt/6 lam after the swap), so only the first step needs
    // adj_start_kernel.
    auto reverse = [&](int s0, int s1, double* Y) -> hipError_t {
        hipError_t er;
        for (int i = s1 - 1; i >= s0; --i) {
            double* Yi = Y + nn * 4 * (size_t)(i - s0);
            if (i == steps - 1)
                hipLaunchKernelGGL(adj_start_kernel, g, b, 0, s, lam, lam2, av0, n, dt / 6.0);
            const AdjointStageFuse f4{lam2, av1, lam, dt / 3.0, dt};
            if ((er = enqueue_vjp(p, Yi + 3 * nn, av0, tab, mu, gth, part, s, &f4)) != hipSuccess)
                return er;
            const AdjointStageFuse f3{lam2, av0, lam, dt / 3.0, 0.5 * dt};
            if ((er = enqueue_vjp(p, Yi + 2 * nn, av1, tab, mu, gth, part, s, &f3)) != hipSuccess)
                return er;
            const AdjointStageFuse f2{lam2, av1, lam, dt / 6.0, 0.5 * dt};
            if ((er = enqueue_vjp(p, Yi + nn, av0, tab, mu, gth, part, s, &f2)) != hipSuccess)
                return er;
            const AdjointStageFuse f1{lam2, nullptr, nullptr, 0.0, 0.0, lam, av0, dt / 6.0};
            if ((er = enqueue_vjp(p, Yi, av1, tab, mu, gth, part, s, &f1)) != hipSuccess)
                return er;
            double* t = lam;
            lam = lam2;
            lam2 = t;
        }
        return hipSuccess;
    };

    if (L.store_all) {
        if ((e = reverse(0, steps, ws + L.seg)) != hipSuccess) return e;
    } else {
        // Segment sg's states go to buffer sg % 2.  The recompute of segment
        // sg - 1 runs on the side stream while the reverse sweep of segment
        // sg runs on s: the reverse sweep of sg waits for its recompute
        // (ev_rec), and the recompute into a buffer waits for the reverse
        // sweep that last read it (ev_rev, recorded after the sweep of
        // sg + 1).  An event's later record does not move an earlier wait.
        double* buf[2] = {ws + L.seg, ws + L.seg2};
        double* rk2 = ws + L.rk2;
        double* mu2 = ws + L.mu2;
        hipStream_t s2 = nullptr;
        if ((e = side_stream(&s2)) != hipSuccess) return e;
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};   // start, rec, rev
        for (int k = 0; k < 3 && e == hipSuccess; ++k)
            e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(ev[0], s);
        if (e == hipSuccess) e = hipStreamWaitEvent(s2, ev[0], 0);
        if (e == hipSuccess) e = enqueue_fill_zero(rk2, n, s2);
        const int last = L.nseg - 1;
        if (e == hipSuccess) e = recompute(last, buf[last % 2], rk2, mu2, s2);
        if (e == hipSuccess) e = hipEventRecord(ev[1], s2);
        for (int sg = last; e == hipSuccess && sg >= 0; --sg) {
            e = hipStreamWaitEvent(s, ev[1], 0);
            if (e == hipSuccess && sg > 0) {
                if (sg < last) e = hipStreamWaitEvent(s2, ev[2], 0);
                if (e == hipSuccess) e = recompute(sg - 1, buf[(sg - 1) % 2], rk2, mu2, s2);
                if (e == hipSuccess) e = hipEventRecord(ev[1], s2);
            }
            const int s0 = sg * every, s1 = min(steps, s0 + every);
            if (e == hipSuccess) e = reverse(s0, s1, buf[sg % 2]);
            if (e == hipSuccess) e = hipEventRecord(ev[2], s);
        }
        for (int k = 0; k < 3; ++k)
            if (ev[k]) (void)hipEventDestroy(ev[k]);
        if (e != hipSuccess) return e;
    }
    if ((e = enqueue_copy(lam, gN0, n, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(gth_reduce_kernel, dim3(1), dim3(256), 0, s, gth, gtheta, n);
    return hipGetLastError();
}

}  // namespace pbe
