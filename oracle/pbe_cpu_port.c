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
 */
#include "pbe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double beta_of(int kernel, double xj, double xk, double b0,
                             double b_br, double b_sh)
{
    return oracle_beta(kernel, xj, xk, b0, b_br, b_sh);
}

int port_aggregation_rhs(int n, const double* N, const double* x, int kernel,
                         double beta0, double beta_br, double beta_sh,
                         double log_x0, double inv_log_r, double* rhs,
                         int threads)
{
    if (n <= 0) return ORACLE_ERR_INVALID_VALUE;
    if (!N || !x || !rhs) return ORACLE_ERR_INVALID_POINTER;
    if (kernel < 0 || kernel > 5) return ORACLE_ERR_INVALID_VALUE;
    if (threads < 1) threads = 1;
    double* priv = (double*)calloc((size_t)n * (size_t)threads, sizeof(double));
    if (!priv) return ORACLE_ERR_NO_MEMORY;
    const double xlast = x[n - 1], xfirst = x[0];

#pragma omp parallel num_threads(threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* r = priv + (size_t)tid * (size_t)n;
#pragma omp for schedule(dynamic, 8)
        for (int j = 0; j < n; ++j) {
            const double xj = x[j], Nj = N[j];
            for (int k = j; k < n; ++k) {
                const double xk = x[k];
                double rate = beta_of(kernel, xj, xk, beta0, beta_br, beta_sh) * Nj * N[k];
                if (j == k) { rate *= 0.5; r[j] -= 2.0 * rate; }
                else        { r[j] -= rate; r[k] -= rate; }
                const double v = xj + xk;
                if (v >= xlast)       r[n - 1] += rate;
                else if (v <= xfirst) r[0] += rate;
                else {
                    int hi;
                    if (inv_log_r != 0.0) {
                        hi = (int)floor((log(v) - log_x0) * inv_log_r) + 1;
                        if (hi < 1) hi = 1;
                        if (hi > n - 1) hi = n - 1;
                    } else {
                        int lo = 0; hi = n - 1;
                        while (hi - lo > 1) {
                            int mid = (lo + hi) / 2;
                            if (x[mid] <= v) lo = mid; else hi = mid;
                        }
                    }
                    const int lo = hi - 1;
                    const double w = (v - x[lo]) / (x[hi] - x[lo]);
                    r[lo] += (1.0 - w) * rate;
                    r[hi] += w * rate;
                }
            }
        }
    }
    for (int t = 0; t < threads; ++t)
        for (int i = 0; i < n; ++i) rhs[i] += priv[(size_t)t * (size_t)n + i];
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
