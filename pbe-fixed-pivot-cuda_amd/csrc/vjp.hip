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
// in a fixed order, so results are bitwise repeatable.  Per-row parameter
// partials are accumulated into a per-row array (each element owned by one
// lane) and reduced in a fixed order at the end.
#include <algorithm>

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
    int sel;                          // -1: no breakage
    double S0, v_ref, alpha, v_min;
    int log_scale;                    // 1: slots 0-3 hold d/d ln(coefficient)
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

// eta(lam) of a near or clipped pair with volume v and larger index K.
__device__ __forceinline__ double near_eta(const VjpArgs& a, double v, int K)
{
    const int n = a.n;
    if (v >= a.x[n - 1]) return a.lam[n - 1];
    if (v <= a.x[0]) return a.lam[0];
    const int hi = bracket_search(a.x, K, n - 1, v);
    const int lo = hi - 1;
    const double xl = a.x[lo], xh = a.x[hi];
    const double w = (v - xl) / (xh - xl);
    return (1.0 - w) * a.lam[lo] + w * a.lam[hi];
}

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

// Rows [m0, m0 + 64) of the instance described by `a`; all 256 threads call
// it.  KT < 0: no aggregation term.
template <int KT>
__device__ __forceinline__ void vjp_rows_body(const VjpArgs& a, const int m0, VjpSmem<KT>& sm)
{
    constexpr int G = VjpSmem<KT>::G;
    auto& part = sm.part;

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n = a.n;
    const int m = m0 + lane;
    const bool row = m < n;

    double acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = 0.0;

    double Nm = 0.0, lm = 0.0, xm = 0.0;
    if (row) {
        Nm = a.N[m];
        lm = a.lam[m];
        xm = a.x[m];
    }

    if constexpr (KT >= 0) {
        using VG = VjpGroups<KT>;
        double xm1 = xm, im = 0.0, lm1 = lm, cm = 0.0, icm = 0.0;
        if (row) {
            cm = a.tab[4 * m + 0];
            icm = a.tab[4 * m + 1];
            im = a.tab[4 * m + 2];
            if (m < n - 1) {
                xm1 = a.x[m + 1];
                lm1 = a.lam[m + 1];
            }
        }
        // this wave's quarter of the columns
        const int chunk = (n + 3) / 4;
        const int kb = wv * chunk, ke = min(n, kb + chunk);
        for (int k = kb; k < ke; ++k) {
            const double xk = a.x[k];
            const double Nk = a.N[k];
            const double lk = a.lam[k];
            const double ck = a.tab[4 * k + 0];
            const double ick = a.tab[4 * k + 1];
            const double ik = a.tab[4 * k + 2];
            const double xk1 = (k < n - 1) ? a.x[k + 1] : xk;
            const double lk1 = (k < n - 1) ? a.lam[k + 1] : lk;
            if (!row) continue;
            const bool kge = k >= m;
            const double xK = kge ? xk : xm;
            const double xK1 = kge ? xk1 : xm1;
            const double iK = kge ? ik : im;
            const double lK = kge ? lk : lm;
            const double lK1 = kge ? lk1 : lm1;
            const double v = xm + xk;
            double eta;
            if (v < xK1) {
                const double w = (v - xK) * iK;
                eta = fma(w, lK1 - lK, lK);
            } else {
                eta = near_eta(a, v, kge ? k : m);
            }
            const double t = Nk * ((eta - lm) - lk);
            double f[G];
            VG::phi(f, xm, cm, icm, xk, ck, ick);
#pragma unroll
            for (int g = 0; g < G; ++g) acc[g] = fma(f[g], t, acc[g]);
        }
    }

    // combine the four column quarters in a fixed order
    if (wv > 0) {
#pragma unroll
        for (int g = 0; g < G; ++g) part[wv - 1][lane][g] = acc[g];
    }
    __syncthreads();
    if (wv != 0 || !row) return;
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = ((acc[g] + part[0][lane][g]) + part[1][lane][g]) + part[2][lane][g];

    double gn = 0.0;
    double* gr = a.gth_rows + (size_t)kNP * m;
    if constexpr (KT >= 0) {
        using VG = VjpGroups<KT>;
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
    a.gN[m] = gn;
}

template <int KT>
__global__ __launch_bounds__(256)
void vjp_kernel(VjpArgs a)
{
    __shared__ VjpSmem<KT> sm;
    vjp_rows_body<KT>(a, blockIdx.x * kWave, sm);
}

// Batched VJP: items are (instance, 64-row block); slots 0-3 of the per-row
// partials hold d/d ln(coefficient) so that their sum over instances is the
// gradient of jointly scaled coefficients.
template <int KT>
__global__ __launch_bounds__(256)
void vjp_batch_kernel(VjpArgs base, const BatchInstDev* __restrict__ inst,
                      const BatchItem* __restrict__ items, int nitems)
{
    __shared__ VjpSmem<KT> sm;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BatchItem w = items[it];
        const BatchInstDev d = inst[w.inst];
        VjpArgs a = base;
        a.N = base.N + d.off;
        a.x = base.x + d.off;
        a.lam = base.lam + d.off;
        a.tab = base.tab + 4 * (size_t)d.off;
        a.gN = base.gN + d.off;
        a.gth_rows = base.gth_rows + (size_t)kNP * d.off;
        a.n = d.n;
        a.coef = KernelCoef{d.b0, d.b_br, d.b_sh};
        a.S0 = d.S0;
        a.v_ref = d.v_ref;
        a.alpha = d.alpha;
        a.v_min = d.v_min;
        vjp_rows_body<KT>(a, w.unit * kWave, sm);
        __syncthreads();
    }
}

hipError_t enqueue_vjp_table(const double* x, double* tab, int n, hipStream_t s)
{
    hipLaunchKernelGGL(vjp_table_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, tab, n);
    return hipGetLastError();
}

hipError_t enqueue_vjp(const RhsProblem& p, const double* N, const double* lam,
                       const double* tab, double* gN, double* gth_rows, hipStream_t s)
{
    VjpArgs a{};
    a.N = N;
    a.x = p.x;
    a.t_q = p.t_q;
    a.bw_q = p.bw_q;
    a.lam = lam;
    a.tab = tab;
    a.gN = gN;
    a.gth_rows = gth_rows;
    a.n = p.n;
    a.nq = p.brk ? p.brk->n_quad : 0;
    a.sel = p.brk ? p.brk->selection : -1;
    if (p.agg) a.coef = KernelCoef{p.agg->beta0, p.agg->beta_br, p.agg->beta_sh};
    if (p.brk) {
        a.S0 = p.brk->S0;
        a.v_ref = p.brk->v_ref;
        a.alpha = p.brk->alpha;
        a.v_min = p.brk->v_min;
    }
    if (p.batch) {
        const pbe_batch_plan_s* pl = p.batch;
        a.nq = pl->selection >= 0 ? pl->n_quad : 0;
        a.sel = pl->selection;
        a.log_scale = 1;
        const dim3 g((unsigned)std::min(pl->grid, pl->n_row_items)), b(256);
        const BatchInstDev* di = pl->d_inst;
        const BatchItem* dr = pl->d_row_items;
        const int ni = pl->n_row_items;
        switch (pl->kernel_type) {
        case -1: hipLaunchKernelGGL(vjp_batch_kernel<-1>, g, b, 0, s, a, di, dr, ni); break;
        case 0: hipLaunchKernelGGL(vjp_batch_kernel<0>, g, b, 0, s, a, di, dr, ni); break;
        case 1: hipLaunchKernelGGL(vjp_batch_kernel<1>, g, b, 0, s, a, di, dr, ni); break;
        case 2: hipLaunchKernelGGL(vjp_batch_kernel<2>, g, b, 0, s, a, di, dr, ni); break;
        case 3: hipLaunchKernelGGL(vjp_batch_kernel<3>, g, b, 0, s, a, di, dr, ni); break;
        case 4: hipLaunchKernelGGL(vjp_batch_kernel<4>, g, b, 0, s, a, di, dr, ni); break;
        case 5: hipLaunchKernelGGL(vjp_batch_kernel<5>, g, b, 0, s, a, di, dr, ni); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const dim3 g((unsigned)((p.n + kWave - 1) / kWave)), b(256);
    switch (p.agg ? p.agg->kernel_type : -1) {
    case -1: hipLaunchKernelGGL(vjp_kernel<-1>, g, b, 0, s, a); break;
    case 0: hipLaunchKernelGGL(vjp_kernel<0>, g, b, 0, s, a); break;
    case 1: hipLaunchKernelGGL(vjp_kernel<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(vjp_kernel<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(vjp_kernel<3>, g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL(vjp_kernel<4>, g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL(vjp_kernel<5>, g, b, 0, s, a); break;
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

// lam_new += mu; av = c1 lam + c2 mu (next stage's cotangent)
__global__ void adj_stage_kernel(const double* __restrict__ lam, const double* __restrict__ mu,
                                 double* __restrict__ lam_new, double* __restrict__ av, int n,
                                 double c1, double c2, int last)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mu[i];
    lam_new[i] += m;
    if (!last) av[i] = fma(c2, m, c1 * lam[i]);
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

AdjointLayout adjoint_layout(int n, int steps, int every)
{
    AdjointLayout L{};
    const size_t nn = (size_t)n;
    L.nseg = (steps + every - 1) / every;
    size_t off = 0;
    L.ckpt = off;  off += nn * (size_t)L.nseg;
    L.seg = off;   off += nn * 4 * (size_t)every;
    L.tab = off;   off += 4 * nn;
    L.lam = off;   off += nn;
    L.lam2 = off;  off += nn;
    L.av = off;    off += nn;
    L.mu = off;    off += nn;
    L.rk = off;    off += 3 * nn;
    L.gth = off;   off += (size_t)kNP * nn;
    L.total = off; // doubles
    return L;
}

hipError_t enqueue_adjoint_forward(const RhsProblem& p, double* N, double dt, int steps,
                                   int every, double* ws, hipStream_t s)
{
    const AdjointLayout L = adjoint_layout(p.n, steps, every);
    double* rk = ws + L.rk;
    hipError_t e = enqueue_fill_zero(rk, p.n, s);
    for (int i = 0; e == hipSuccess && i < steps; ++i) {
        if (i % every == 0) e = enqueue_copy(N, ws + L.ckpt + (size_t)p.n * (i / every), p.n, s);
        if (e == hipSuccess) e = enqueue_rk4_step_stages(p, N, dt, rk, nullptr, s);
    }
    return e;
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
    double* av = ws + L.av;
    double* mu = ws + L.mu;
    double* rk = ws + L.rk;
    double* gth = ws + L.gth;
    double* seg = ws + L.seg;
    hipError_t e;
    if ((e = enqueue_vjp_table(p.x, tab, n, s)) != hipSuccess) return e;
    if ((e = enqueue_fill_zero(gth, kNP * n, s)) != hipSuccess) return e;
    if ((e = enqueue_fill_zero(rk, n, s)) != hipSuccess) return e;
    if ((e = enqueue_copy(gN_final, lam, n, s)) != hipSuccess) return e;
    const dim3 g = grid1(n), b(256);
    for (int sg = L.nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * every, s1 = min(steps, s0 + every);
        // recompute the segment, keeping every stage state
        double* Y = seg;   // Y[(i - s0)*4 + q]
        if ((e = enqueue_copy(ws + L.ckpt + nn * sg, Y, n, s)) != hipSuccess) return e;
        for (int i = s0; i < s1; ++i) {
            double* Yi = Y + nn * 4 * (size_t)(i - s0);
            double* st[3] = {Yi + nn, Yi + 2 * nn, Yi + 3 * nn};
            // the step advances a copy: Y1 of step i+1 follows Y4 of step i
            double* next = (i + 1 < s1) ? Yi + 4 * nn : nullptr;
            if (next) {
                if ((e = enqueue_copy(Yi, next, n, s)) != hipSuccess) return e;
                if ((e = enqueue_rk4_step_stages(p, next, dt, rk, st, s)) != hipSuccess) return e;
            } else {
                // last step of the segment: only the stage states are needed
                if ((e = enqueue_copy(Yi, mu, n, s)) != hipSuccess) return e;
                if ((e = enqueue_rk4_step_stages(p, mu, dt, rk, st, s)) != hipSuccess) return e;
            }
        }
        // reverse sweep over the segment
        for (int i = s1 - 1; i >= s0; --i) {
            double* Yi = Y + nn * 4 * (size_t)(i - s0);
            hipLaunchKernelGGL(adj_start_kernel, g, b, 0, s, lam, lam2, av, n, dt / 6.0);
            if ((e = enqueue_vjp(p, Yi + 3 * nn, av, tab, mu, gth, s)) != hipSuccess) return e;
            hipLaunchKernelGGL(adj_stage_kernel, g, b, 0, s, lam, mu, lam2, av, n, dt / 3.0, dt, 0);
            if ((e = enqueue_vjp(p, Yi + 2 * nn, av, tab, mu, gth, s)) != hipSuccess) return e;
            hipLaunchKernelGGL(adj_stage_kernel, g, b, 0, s, lam, mu, lam2, av, n, dt / 3.0, 0.5 * dt, 0);
            if ((e = enqueue_vjp(p, Yi + nn, av, tab, mu, gth, s)) != hipSuccess) return e;
            hipLaunchKernelGGL(adj_stage_kernel, g, b, 0, s, lam, mu, lam2, av, n, dt / 6.0, 0.5 * dt, 0);
            if ((e = enqueue_vjp(p, Yi, av, tab, mu, gth, s)) != hipSuccess) return e;
            hipLaunchKernelGGL(adj_stage_kernel, g, b, 0, s, lam, mu, lam2, av, n, 0.0, 0.0, 1);
            double* t = lam;
            lam = lam2;
            lam2 = t;
        }
    }
    if ((e = enqueue_copy(lam, gN0, n, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(gth_reduce_kernel, dim3(1), dim3(256), 0, s, gth, gtheta, n);
    return hipGetLastError();
}

}  // namespace pbe
