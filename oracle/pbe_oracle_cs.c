/*
 * pbe_oracle_cs.c -- complex-step restatement of the fixed-pivot right-hand
 * side, explicit RK4 and a log-PSD misfit loss, for gradient checks.
 * TEST INFRASTRUCTURE ONLY (the checker of the adjoint; never part of the
 * product library).
 *
 * The arithmetic follows src/aggregation_kernels.cuh:52-89,186-231 and
 * src/breakage_kernels.cuh:51-65,126-242 of the reference and the RK4 of
 * examples/common/pbe_examples_utils.hpp:416-457, with N and the parameters
 * (beta0, beta_br, beta_sh, S0, alpha) complex: evaluating at theta + i*h*e
 * gives dF/dtheta = Im F / h to machine precision (complex-step
 * differentiation), because nothing in the map is non-analytic in them --
 * brackets and clip tests depend on the real pivots only, and the breakage
 * guard (N > 2.2e-308) and the loss floor are tested on real parts.
 * The bracket is the exact one (x[lo] <= v < x[hi]).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

typedef struct {
    int n;
    const double* x;
    int use_agg, kernel;
    int use_brk, nq;
    const double* t_q;
    const double* bw_q;
    int sel;
    double v_ref, v_min;
    int threads;
} cs_problem;

enum { TH_B0 = 0, TH_BBR = 1, TH_BSH = 2, TH_S0 = 3, TH_ALPHA = 4 };

static cplx beta_cs(int kernel, double xj, double xk, const cplx* th)
{
    switch (kernel) {
    case 0: return th[TH_B0];
    case 1: return th[TH_B0] * (xj + xk);
    case 2: return th[TH_B0] * xj * xk;
    case 3: {
        double a = cbrt(xj), b = cbrt(xk);
        return th[TH_BBR] * (a / b + b / a + 2.0);
    }
    case 4: {
        double s = cbrt(xj) + cbrt(xk);
        return th[TH_BSH] * (s * s * s);
    }
    default: {
        double a = cbrt(xj), b = cbrt(xk);
        double s = a + b;
        return th[TH_BBR] * (a / b + b / a + 2.0) + th[TH_BSH] * (s * s * s);
    }
    }
}

static cplx selection_cs(int sel, double v, const cplx* th, double v_ref, double v_min)
{
    switch (sel) {
    case 0: return th[TH_S0];
    case 1: return th[TH_S0] * v / v_ref;
    case 2: return th[TH_S0] * cexp(th[TH_ALPHA] * log(v / v_ref));
    default: return (v > v_min) ? th[TH_S0] : 0.0;
    }
}

static int hi_exact(const double* x, int n, double v)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    return hi;
}

static void rhs_cs(const cs_problem* p, const cplx* th, const cplx* N, cplx* out,
                   cplx* priv)
{
    const int n = p->n;
    const double* x = p->x;
    int T = p->threads > 0 ? p->threads : 1;
    memset(priv, 0, sizeof(cplx) * (size_t)n * (size_t)T);
#pragma omp parallel num_threads(T)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        cplx* r = priv + (size_t)tid * (size_t)n;
        if (p->use_agg) {
#pragma omp for schedule(dynamic, 4)
            for (int j = 0; j < n; ++j) {
                for (int k = j; k < n; ++k) {
                    cplx rate = beta_cs(p->kernel, x[j], x[k], th) * N[j] * N[k];
                    if (j == k) { rate *= 0.5; r[j] -= 2.0 * rate; }
                    else { r[j] -= rate; r[k] -= rate; }
                    double v = x[j] + x[k];
                    if (v >= x[n - 1]) r[n - 1] += rate;
                    else if (v <= x[0]) r[0] += rate;
                    else {
                        int hi = hi_exact(x, n, v), lo = hi - 1;
                        double w = (v - x[lo]) / (x[hi] - x[lo]);
                        r[lo] += (1.0 - w) * rate;
                        r[hi] += w * rate;
                    }
                }
            }
        }
        if (p->use_brk) {
#pragma omp for schedule(static)
            for (int j = 0; j < n; ++j) {
                cplx S = selection_cs(p->sel, x[j], th, p->v_ref, p->v_min);
                if (!(creal(S) != 0.0 && creal(N[j]) > 2.2e-308)) continue;
                cplx sn = S * N[j];
                r[j] -= sn;
                for (int q = 0; q < p->nq; ++q) {
                    cplx c = sn * p->bw_q[q];
                    double v = x[j] * p->t_q[q];
                    if (v <= x[0]) r[0] += c;
                    else if (v >= x[n - 1]) r[n - 1] += c;
                    else {
                        int hi = hi_exact(x, n, v), lo = hi - 1;
                        double e = (v - x[lo]) / (x[hi] - x[lo]);
                        r[lo] += (1.0 - e) * c;
                        r[hi] += e * c;
                    }
                }
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        cplx s = 0.0;
        for (int t = 0; t < T; ++t) s += priv[(size_t)t * (size_t)n + i];
        out[i] = s;
    }
}

static void load_theta(const double* re, const double* im, cplx* th)
{
    for (int i = 0; i < 5; ++i) th[i] = re[i] + I * (im ? im[i] : 0.0);
}

/* out = R(N) for complex N and parameters; arrays split into re/im parts */
int cs_rhs(const cs_problem* p, const double* th_re, const double* th_im,
           const double* Nre, const double* Nim, double* ore, double* oim)
{
    const int n = p->n;
    int T = p->threads > 0 ? p->threads : 1;
    cplx th[5];
    load_theta(th_re, th_im, th);
    cplx* N = malloc(sizeof(cplx) * (size_t)n);
    cplx* o = malloc(sizeof(cplx) * (size_t)n);
    cplx* priv = malloc(sizeof(cplx) * (size_t)n * (size_t)T);
    if (!N || !o || !priv) { free(N); free(o); free(priv); return 3; }
    for (int i = 0; i < n; ++i) N[i] = Nre[i] + I * (Nim ? Nim[i] : 0.0);
    rhs_cs(p, th, N, o, priv);
    for (int i = 0; i < n; ++i) { ore[i] = creal(o[i]); oim[i] = cimag(o[i]); }
    free(N); free(o); free(priv);
    return 0;
}

/* RK4 for `steps` steps from complex N0, then
 * L = sum_i mask_i (l_i - logT_i)^2 with l_i = log(N_i) if Re N_i > eps,
 * log(eps) otherwise.  Returns L (re, im) and the final state. */
int cs_rk4_loss(const cs_problem* p, const double* th_re, const double* th_im,
                const double* N0re, const double* N0im, double dt, int steps,
                const double* logT, const double* mask, double eps,
                double* L_re, double* L_im, double* Nf_re, double* Nf_im)
{
    const int n = p->n;
    int T = p->threads > 0 ? p->threads : 1;
    cplx th[5];
    load_theta(th_re, th_im, th);
    cplx* buf = malloc(sizeof(cplx) * (size_t)n * 6);
    cplx* priv = malloc(sizeof(cplx) * (size_t)n * (size_t)T);
    if (!buf || !priv) { free(buf); free(priv); return 3; }
    cplx *N = buf, *k1 = buf + n, *k2 = buf + 2 * n, *k3 = buf + 3 * n,
         *k4 = buf + 4 * n, *st = buf + 5 * n;
    for (int i = 0; i < n; ++i) N[i] = N0re[i] + I * (N0im ? N0im[i] : 0.0);
    for (int s = 0; s < steps; ++s) {
        rhs_cs(p, th, N, k1, priv);
        for (int i = 0; i < n; ++i) st[i] = N[i] + 0.5 * dt * k1[i];
        rhs_cs(p, th, st, k2, priv);
        for (int i = 0; i < n; ++i) st[i] = N[i] + 0.5 * dt * k2[i];
        rhs_cs(p, th, st, k3, priv);
        for (int i = 0; i < n; ++i) st[i] = N[i] + dt * k3[i];
        rhs_cs(p, th, st, k4, priv);
        for (int i = 0; i < n; ++i)
            N[i] += (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    }
    cplx L = 0.0;
    for (int i = 0; i < n; ++i) {
        if (mask[i] == 0.0) continue;
        cplx l = (creal(N[i]) > eps) ? clog(N[i]) : log(eps);
        cplx d = l - logT[i];
        L += mask[i] * d * d;
    }
    *L_re = creal(L);
    *L_im = cimag(L);
    if (Nf_re) for (int i = 0; i < n; ++i) { Nf_re[i] = creal(N[i]); Nf_im[i] = cimag(N[i]); }
    free(buf);
    free(priv);
    return 0;
}
