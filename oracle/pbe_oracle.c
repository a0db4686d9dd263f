/*
 * pbe_oracle.c -- CPU restatement of the fixed-pivot aggregation and breakage
 * right-hand sides of acuoci/pbe-fixed-pivot-cuda.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels are
 * compared against; it is never linked into the product library and the
 * product never calls it.
 *
 * Semantics followed, line by line (paths relative to the reference repo):
 *   beta formulas ............ src/aggregation_kernels.cuh:52-89
 *   bin lookup (both paths) .. src/aggregation_kernels.cuh:118-134
 *   pair rate / clip / split . src/aggregation_kernels.cuh:186-200, 212-231
 *   selection functions ...... src/breakage_kernels.cuh:51-65
 *   left-bracket search ...... src/breakage_kernels.cuh:95-105
 *   death guard .............. src/breakage_kernels.cuh:126-132
 *   birth guard/clip/split ... src/breakage_kernels.cuh:189-215, 231-242
 *   argument validation ...... src/aggregation.cu:23-24,54 ; src/breakage.cu:23-24
 *   RK4 arithmetic order ..... examples/common/pbe_examples_utils.hpp:416-457
 *
 * Every contribution is formed with exactly the reference's floating-point
 * expression and then accumulated per bin in double-double (TwoSum), so the
 * result is the reference's set of contributions summed essentially exactly,
 * independent of order.  Alongside rhs the oracle returns, per bin, the sum of
 * absolute contributions A_i (the scale of the per-entry parity metric) and the
 * exact first-moment defect sum_i x_i rhs_i implied by boundary clipping and
 * by the daughter table's volume moment.
 *
 * Bracket modes for aggregation:
 *   ORACLE_BRACKET_REFERENCE: the reference's own rule -- log-ratio formula
 *       when inv_log_r != 0, binary search otherwise.
 *   ORACLE_BRACKET_EXACT: the unique bracket x[lo] <= v < x[hi] by comparison
 *       (what the reference's binary search returns, and what the log formula
 *       returns whenever log() rounding does not move floor() across an
 *       integer).
 */
#include "pbe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* double-double accumulator (Knuth TwoSum)                                 */
/* ------------------------------------------------------------------------ */
typedef struct { double hi, lo; } dd_t;

static inline void dd_add(dd_t* a, double b)
{
    double s  = a->hi + b;
    double bp = s - a->hi;
    double e  = (a->hi - (s - bp)) + (b - bp);
    a->hi = s;
    a->lo += e;
}

static inline double dd_value(const dd_t* a) { return a->hi + a->lo; }

typedef struct {
    dd_t*   acc;     /* per-bin contribution sum */
    double* abs;     /* per-bin sum of |contribution| (may be NULL) */
} sink_t;

static inline void sink_add(sink_t* s, int i, double c)
{
    dd_add(&s->acc[i], c);
    if (s->abs) s->abs[i] += fabs(c);
}

/* ------------------------------------------------------------------------ */
/* beta(x_j, x_k): src/aggregation_kernels.cuh:52-89, same operation order  */
/* ------------------------------------------------------------------------ */
double oracle_beta(int kernel, double xj, double xk,
                   double b0, double b_br, double b_sh)
{
    switch (kernel) {
    case 0: return b0;
    case 1: return b0 * (xj + xk);
    case 2: return b0 * xj * xk;
    case 3: {
        double xj3 = cbrt(xj), xk3 = cbrt(xk);
        return b_br * (xj3 / xk3 + xk3 / xj3 + 2.0);
    }
    case 4: {
        double s = cbrt(xj) + cbrt(xk);
        return b_sh * s * s * s;
    }
    case 5: {
        double xj3 = cbrt(xj), xk3 = cbrt(xk);
        double br  = b_br * (xj3 / xk3 + xk3 / xj3 + 2.0);
        double sh  = b_sh * (xj3 + xk3) * (xj3 + xk3) * (xj3 + xk3);
        return br + sh;
    }
    default: return 0.0;
    }
}

/* Right bracket by the reference's rule: src/aggregation_kernels.cuh:118-134 */
static int bin_index_reference(const double* x, int n, double v,
                               double log_x0, double inv_log_r)
{
    if (inv_log_r != 0.0) {
        double pos = (log(v) - log_x0) * inv_log_r;
        int hi = (int)floor(pos) + 1;
        if (hi < 1) hi = 1;
        if (hi > n - 1) hi = n - 1;
        return hi;
    }
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    return hi;
}

/* Unique right bracket x[hi-1] <= v < x[hi] (precondition x[0] < v < x[n-1]) */
static int bin_index_exact(const double* x, int n, double v)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    return hi;
}

/* The pairs (j, k >= j) of rows [j_lo, j_hi) into sink s and the clip
 * defect d: src/aggregation_kernels.cuh:186-200 and :212-231. */
static void agg_rows(int n, const double* N, const double* x, int kernel, double beta0,
                     double beta_br, double beta_sh, double log_x0, double inv_log_r,
                     int bracket_mode, int j_lo, int j_hi, sink_t* s, dd_t* defect)
{
    for (int j = j_lo; j < j_hi; ++j) {
        for (int k = j; k < n; ++k) {
            const int diagonal = (j == k);
            double xj = x[j], xk = x[k], Nj = N[j], Nk = N[k];
            /* :186-188 */
            double beta = oracle_beta(kernel, xj, xk, beta0, beta_br, beta_sh);
            double rate = beta * Nj * Nk;
            if (diagonal) rate *= 0.5;
            /* death, :213-221 */
            if (diagonal) {
                sink_add(s, j, -2.0 * rate);
            } else {
                sink_add(s, j, -rate);
                sink_add(s, k, -rate);
            }
            /* birth, :190-200 and :223-231 */
            double v = xj + xk;
            if (v >= x[n - 1]) {
                sink_add(s, n - 1, rate);
                dd_add(defect, rate * (x[n - 1] - v));
            } else if (v <= x[0]) {
                sink_add(s, 0, rate);
                dd_add(defect, rate * (x[0] - v));
            } else {
                int hi = (bracket_mode == ORACLE_BRACKET_REFERENCE)
                         ? bin_index_reference(x, n, v, log_x0, inv_log_r)
                         : bin_index_exact(x, n, v);
                int lo = hi - 1;
                double w = (v - x[lo]) / (x[hi] - x[lo]);
                sink_add(s, lo, (1.0 - w) * rate);
                sink_add(s, hi, w * rate);
            }
        }
    }
}

int oracle_aggregation_rhs_threads(int n, const double* N, const double* x,
                                   int kernel, double beta0, double beta_br,
                                   double beta_sh, double log_x0, double inv_log_r,
                                   int bracket_mode, double* rhs, double* abs_sum,
                                   double* m1_defect, int threads)
{
    /* validation order of src/aggregation.cu:23-24 and :54 */
    if (n <= 0) return ORACLE_ERR_INVALID_VALUE;
    if (!N || !x || !rhs) return ORACLE_ERR_INVALID_POINTER;
    if (kernel < 0 || kernel > 5) return ORACLE_ERR_INVALID_VALUE;
    if (threads < 1) threads = 1;
    if (threads > n) threads = n;

    /* Row ranges of about equal pair counts, one per thread; each thread
     * accumulates into its own double-double bins and |contribution| sums,
     * merged afterwards in thread order.  Double-double sums are exact to
     * ~2^-106 relative, so the result does not depend on the thread count
     * beyond that. */
    dd_t* acc = (dd_t*)calloc((size_t)n * (size_t)threads, sizeof(dd_t));
    double* abs_t = abs_sum ? (double*)calloc((size_t)n * (size_t)threads, sizeof(double))
                            : NULL;
    dd_t* defect_t = (dd_t*)calloc((size_t)threads, sizeof(dd_t));
    int* bound = (int*)malloc(sizeof(int) * (size_t)(threads + 1));
    if (!acc || (abs_sum && !abs_t) || !defect_t || !bound) {
        free(acc); free(abs_t); free(defect_t); free(bound);
        return ORACLE_ERR_NO_MEMORY;
    }
    /* rows j >= b hold (n - b)(n - b + 1)/2 pairs */
    const double total = 0.5 * (double)n * (double)(n + 1);
    bound[0] = 0;
    for (int t = 1; t < threads; ++t) {
        const double rest = total * (double)(threads - t) / (double)threads;
        int b = n - (int)floor(sqrt(2.0 * rest));
        if (b < bound[t - 1]) b = bound[t - 1];
        if (b > n) b = n;
        bound[t] = b;
    }
    bound[threads] = n;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (int t = 0; t < threads; ++t) {
        sink_t s;
        s.acc = acc + (size_t)t * (size_t)n;
        s.abs = abs_t ? abs_t + (size_t)t * (size_t)n : NULL;
        agg_rows(n, N, x, kernel, beta0, beta_br, beta_sh, log_x0, inv_log_r, bracket_mode,
                 bound[t], bound[t + 1], &s, &defect_t[t]);
    }
    dd_t defect = {0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        dd_t a = {0.0, 0.0};
        double ab = 0.0;
        for (int t = 0; t < threads; ++t) {
            const dd_t* p = &acc[(size_t)t * (size_t)n + i];
            dd_add(&a, p->hi);
            dd_add(&a, p->lo);
            if (abs_t) ab += abs_t[(size_t)t * (size_t)n + i];
        }
        rhs[i] += dd_value(&a);
        if (abs_sum) abs_sum[i] += ab;
    }
    for (int t = 0; t < threads; ++t) {
        dd_add(&defect, defect_t[t].hi);
        dd_add(&defect, defect_t[t].lo);
    }
    if (m1_defect) *m1_defect = dd_value(&defect);
    free(acc); free(abs_t); free(defect_t); free(bound);
    return 0;
}

int oracle_aggregation_rhs(int n, const double* N, const double* x,
                           int kernel, double beta0, double beta_br,
                           double beta_sh, double log_x0, double inv_log_r,
                           int bracket_mode, double* rhs, double* abs_sum,
                           double* m1_defect)
{
    return oracle_aggregation_rhs_threads(n, N, x, kernel, beta0, beta_br, beta_sh, log_x0,
                                          inv_log_r, bracket_mode, rhs, abs_sum, m1_defect, 1);
}

/* ------------------------------------------------------------------------ */
/* selection S(v): src/breakage_kernels.cuh:51-65                            */
/* ------------------------------------------------------------------------ */
double oracle_selection(int sel, double v, double S0, double v_ref,
                        double alpha, double v_min)
{
    switch (sel) {
    case 0: return S0;
    case 1: return S0 * v / v_ref;
    case 2: return S0 * pow(v / v_ref, alpha);
    case 3: return (v > v_min) ? S0 : 0.0;
    default: return 0.0;
    }
}

/* left bracket: src/breakage_kernels.cuh:95-105 */
static int bin_search_left(const double* x, int n, double v)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

int oracle_breakage_rhs(int n, int n_quad, const double* N, const double* x,
                        const double* t_q, const double* bw_q, int selection,
                        double S0, double v_ref, double alpha, double v_min,
                        double* rhs, double* abs_sum, double* m1_defect)
{
    /* src/breakage.cu:23-24, then the enum check of :46/:74 */
    if (n <= 0 || n_quad <= 0) return ORACLE_ERR_INVALID_VALUE;
    if (!N || !x || !t_q || !bw_q || !rhs) return ORACLE_ERR_INVALID_POINTER;
    if (selection < 0 || selection > 3) return ORACLE_ERR_INVALID_VALUE;

    sink_t s;
    s.acc = (dd_t*)calloc((size_t)n, sizeof(dd_t));
    s.abs = abs_sum;
    if (!s.acc) return ORACLE_ERR_NO_MEMORY;
    dd_t defect = {0.0, 0.0};

    for (int j = 0; j < n; ++j) {
        double xj = x[j], Nj = N[j];
        double Sj = oracle_selection(selection, xj, S0, v_ref, alpha, v_min);
        /* guard of :131 and :195 */
        if (!(Sj != 0.0 && Nj > 2.2e-308)) continue;
        /* death, :132 */
        sink_add(&s, j, -(Sj * Nj));
        dd_add(&defect, -(Sj * Nj) * xj);
        for (int q = 0; q < n_quad; ++q) {
            /* :197-215 */
            double contrib = Sj * Nj * bw_q[q];
            double v = xj * t_q[q];
            if (v <= x[0]) {
                sink_add(&s, 0, contrib);
                dd_add(&defect, contrib * x[0]);
            } else if (v >= x[n - 1]) {
                sink_add(&s, n - 1, contrib);
                dd_add(&defect, contrib * x[n - 1]);
            } else {
                int lo = bin_search_left(x, n, v);
                int hi = lo + 1;
                double eta = (v - x[lo]) / (x[hi] - x[lo]);
                sink_add(&s, lo, (1.0 - eta) * contrib);
                sink_add(&s, hi, eta * contrib);
                dd_add(&defect, contrib * v);
            }
        }
    }
    for (int i = 0; i < n; ++i) rhs[i] += dd_value(&s.acc[i]);
    if (m1_defect) *m1_defect = dd_value(&defect);
    free(s.acc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* RK4 driver: examples/common/pbe_examples_utils.hpp:416-457 arithmetic;   */
/* the rhs of each stage is zeroed, then aggregation, then breakage         */
/* (examples/flocculation_two_stage/main.cpp:333-351).                      */
/* ------------------------------------------------------------------------ */
static int eval_rhs(const oracle_problem_t* p, const double* Nin, double* k)
{
    memset(k, 0, sizeof(double) * (size_t)p->n);
    int e = 0;
    if (p->use_agg)
        e = oracle_aggregation_rhs_threads(p->n, Nin, p->x, p->kernel, p->beta0,
                                           p->beta_br, p->beta_sh, p->log_x0,
                                           p->inv_log_r, p->bracket_mode, k, NULL, NULL,
                                           p->threads);
    if (e) return e;
    if (p->use_brk)
        e = oracle_breakage_rhs(p->n, p->n_quad, Nin, p->x, p->t_q, p->bw_q,
                                p->selection, p->S0, p->v_ref, p->alpha,
                                p->v_min, k, NULL, NULL);
    return e;
}

int oracle_rk4(const oracle_problem_t* p, double* Nst, double dt, int steps)
{
    int n = p->n;
    double* buf = (double*)malloc(sizeof(double) * (size_t)n * 5);
    if (!buf) return ORACLE_ERR_NO_MEMORY;
    double *k1 = buf, *k2 = buf + n, *k3 = buf + 2 * n, *k4 = buf + 3 * n,
           *st = buf + 4 * n;
    int e = 0;
    for (int step = 0; step < steps && !e; ++step) {
        if ((e = eval_rhs(p, Nst, k1))) break;
        for (int i = 0; i < n; ++i) st[i] = Nst[i] + 0.5 * dt * k1[i];
        if ((e = eval_rhs(p, st, k2))) break;
        for (int i = 0; i < n; ++i) st[i] = Nst[i] + 0.5 * dt * k2[i];
        if ((e = eval_rhs(p, st, k3))) break;
        for (int i = 0; i < n; ++i) st[i] = Nst[i] + dt * k3[i];
        if ((e = eval_rhs(p, st, k4))) break;
        for (int i = 0; i < n; ++i)
            Nst[i] += (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    }
    free(buf);
    return e;
}
