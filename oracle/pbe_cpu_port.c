/*
 * pbe_cpu_port.c -- plain FP64 OpenMP restatement of the reference's
 * aggregation and breakage right-hand sides, used only as the CPU timing
 * baseline (TEST INFRASTRUCTURE; never part of the product library).
 *
 * Same arithmetic as src/aggregation_kernels.cuh:186-231 and
 * src/breakage_kernels.cuh:126-242 of the reference, with the reference's own
 * bracket rule (log-ratio formula when inv_log_r != 0, binary search
 * otherwise).  Work split: rows j of the upper triangle dealt out dynamically
 * to threads, each thread accumulates into a private rhs, then a reduction.
 * Built at -O3 with the collision kernel inlined per kernel type and the
 * row loops cloned for AVX-512 / AVX2 / baseline, picked at load time for
 * the host that runs it.
 */
#include "pbe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* beta(x_j, x_k) of src/aggregation_kernels.cuh:45-83 with the bins' cube
 * roots taken from a per-call table (cbrt is a pure function of x, so the
 * values are the reference's).  `kernel` is a compile-time constant at every
 * call site below, so the switch folds away and the expression inlines. */
static inline __attribute__((always_inline)) double port_beta(
    const int kernel, double xj, double xk, double cj, double ck, double b0,
    double b_br, double b_sh)
{
    switch (kernel) {
    case 0: return b0;
    case 1: return b0 * (xj + xk);
    case 2: return b0 * xj * xk;
    case 3: return b_br * (cj / ck + ck / cj + 2.0);
    case 4: { const double s = cj + ck; return b_sh * s * s * s; }
    case 5: {
        const double s = cj + ck;
        return b_br * (cj / ck + ck / cj + 2.0) + b_sh * s * s * s;
    }
    default: return 0.0;
    }
}

/* One row j of the upper triangle into the private rhs r: rates of the row
 * first (a loop the compiler vectorises), then deaths, then births with the
 * reference's bracket rule (log-ratio formula when inv_log_r != 0, binary
 * search otherwise; src/aggregation_kernels.cuh:118-134,186-231).  The
 * clones pick AVX-512, AVX2 or baseline code for the host that runs it. */
#define PORT_AGG_ROW(K)                                                                   \
__attribute__((target_clones("avx512f", "avx2", "default")))                              \
static void port_agg_row_##K(int n, int j, const double* N, const double* x,               \
                             const double* cb, double b0, double b_br, double b_sh,        \
                             double log_x0, double inv_log_r, double* rate, double* r)    \
{                                                                                          \
    const double xj = x[j], Nj = N[j], cj = cb[j];                                         \
    const double xlast = x[n - 1], xfirst = x[0];                                          \
    for (int k = j; k < n; ++k)                                                            \
        rate[k] = port_beta(K, xj, x[k], cj, cb[k], b0, b_br, b_sh) * Nj * N[k];           \
    rate[j] *= 0.5;                                                                        \
    double dj = 2.0 * rate[j];                                                             \
    for (int k = j + 1; k < n; ++k) {                                                      \
        dj += rate[k];                                                                     \
        r[k] -= rate[k];                                                                   \
    }                                                                                      \
    r[j] -= dj;                                                                            \
    for (int k = j; k < n; ++k) {                                                          \
        const double v = xj + x[k], c = rate[k];                                           \
        if (v >= xlast) { r[n - 1] += c; continue; }                                       \
        if (v <= xfirst) { r[0] += c; continue; }                                          \
        int hi;                                                                            \
        if (inv_log_r != 0.0) {                                                            \
            hi = (int)floor((log(v) - log_x0) * inv_log_r) + 1;                            \
            if (hi < 1) hi = 1;                                                            \
            if (hi > n - 1) hi = n - 1;                                                    \
        } else {                                                                           \
            int lo = 0;                                                                    \
            hi = n - 1;                                                                    \
            while (hi - lo > 1) {                                                          \
                const int mid = (lo + hi) / 2;                                             \
                if (x[mid] <= v) lo = mid; else hi = mid;                                  \
            }                                                                              \
        }                                                                                  \
        const int lo = hi - 1;                                                             \
        const double w = (v - x[lo]) / (x[hi] - x[lo]);                                    \
        r[lo] += (1.0 - w) * c;                                                            \
        r[hi] += w * c;                                                                    \
    }                                                                                      \
}

PORT_AGG_ROW(0)
PORT_AGG_ROW(1)
PORT_AGG_ROW(2)
PORT_AGG_ROW(3)
PORT_AGG_ROW(4)
PORT_AGG_ROW(5)

int port_aggregation_rhs(int n, const double* N, const double* x, int kernel,
                         double beta0, double beta_br, double beta_sh,
                         double log_x0, double inv_log_r, double* rhs,
                         int threads)
{
    if (n <= 0) return ORACLE_ERR_INVALID_VALUE;
    if (!N || !x || !rhs) return ORACLE_ERR_INVALID_POINTER;
    if (kernel < 0 || kernel > 5) return ORACLE_ERR_INVALID_VALUE;
    if (threads < 1) threads = 1;
    /* per thread: a private rhs and a row of rates; shared: cube roots */
    double* priv = (double*)calloc((size_t)2 * (size_t)n * (size_t)threads + (size_t)n,
                                   sizeof(double));
    if (!priv) return ORACLE_ERR_NO_MEMORY;
    double* cb = priv + (size_t)2 * (size_t)n * (size_t)threads;
    for (int i = 0; i < n; ++i) cb[i] = cbrt(x[i]);

#pragma omp parallel num_threads(threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* r = priv + (size_t)2 * (size_t)tid * (size_t)n;
        double* rate = r + n;
#pragma omp for schedule(dynamic, 8)
        for (int j = 0; j < n; ++j) {
            switch (kernel) {
            case 0: port_agg_row_0(n, j, N, x, cb, beta0, beta_br, beta_sh, log_x0, inv_log_r, rate, r); break;
            case 1: port_agg_row_1(n, j, N, x, cb, beta0, beta_br, beta_sh, log_x0, inv_log_r, rate, r); break;
            case 2: port_agg_row_2(n, j, N, x, cb, beta0, beta_br, beta_sh, log_x0, inv_log_r, rate, r); break;
            case 3: port_agg_row_3(n, j, N, x, cb, beta0, beta_br, beta_sh, log_x0, inv_log_r, rate, r); break;
            case 4: port_agg_row_4(n, j, N, x, cb, beta0, beta_br, beta_sh, log_x0, inv_log_r, rate, r); break;
            default: port_agg_row_5(n, j, N, x, cb, beta0, beta_br, beta_sh, log_x0, inv_log_r, rate, r); break;
            }
        }
    }
    for (int t = 0; t < threads; ++t)
        for (int i = 0; i < n; ++i) rhs[i] += priv[(size_t)2 * (size_t)t * (size_t)n + i];
    free(priv);
    return 0;
}

int port_breakage_rhs(int n, int n_quad, const double* N, const double* x,
                      const double* t_q, const double* bw_q, int selection,
                      double S0, double v_ref, double alpha, double v_min,
                      double* rhs, int threads)
{
    if (n <= 0 || n_quad <= 0) return ORACLE_ERR_INVALID_VALUE;
    if (!N || !x || !t_q || !bw_q || !rhs) return ORACLE_ERR_INVALID_POINTER;
    if (selection < 0 || selection > 3) return ORACLE_ERR_INVALID_VALUE;
    if (threads < 1) threads = 1;
    double* priv = (double*)calloc((size_t)n * (size_t)threads, sizeof(double));
    if (!priv) return ORACLE_ERR_NO_MEMORY;

#pragma omp parallel num_threads(threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* r = priv + (size_t)tid * (size_t)n;
#pragma omp for schedule(static)
        for (int j = 0; j < n; ++j) {
            const double xj = x[j], Nj = N[j];
            const double Sj = oracle_selection(selection, xj, S0, v_ref, alpha, v_min);
            if (!(Sj != 0.0 && Nj > 2.2e-308)) continue;
            r[j] -= Sj * Nj;
            for (int q = 0; q < n_quad; ++q) {
                const double c = Sj * Nj * bw_q[q];
                const double v = xj * t_q[q];
                if (v <= x[0]) r[0] += c;
                else if (v >= x[n - 1]) r[n - 1] += c;
                else {
                    int lo = 0, hi = n - 1;
                    while (hi - lo > 1) {
                        int mid = (lo + hi) >> 1;
                        if (x[mid] <= v) lo = mid; else hi = mid;
                    }
                    hi = lo + 1;
                    const double eta = (v - x[lo]) / (x[hi] - x[lo]);
                    r[lo] += (1.0 - eta) * c;
                    r[hi] += eta * c;
                }
            }
        }
    }
    for (int t = 0; t < threads; ++t)
        for (int i = 0; i < n; ++i) rhs[i] += priv[(size_t)t * (size_t)n + i];
    free(priv);
    return 0;
}
