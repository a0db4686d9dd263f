/*
 * pbe_oracle.h -- CPU restatement of the reference's fixed-pivot source terms.
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's
 * cpu_baseline leg as the checker; never part of the product library.
 */
#ifndef PBE_ORACLE_H
#define PBE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_ERR_INVALID_VALUE   = 1,
    ORACLE_ERR_INVALID_POINTER = 2,
    ORACLE_ERR_NO_MEMORY       = 3
};

enum { ORACLE_BRACKET_REFERENCE = 0, ORACLE_BRACKET_EXACT = 1 };

double oracle_beta(int kernel, double xj, double xk,
                   double b0, double b_br, double b_sh);
double oracle_selection(int sel, double v, double S0, double v_ref,
                        double alpha, double v_min);

/* rhs += aggregation source; abs_sum += |contributions| per bin (may be NULL);
 * *m1_defect = exact sum_i x_i rhs_i of this call (clip defect; may be NULL). */
int oracle_aggregation_rhs(int n, const double* N, const double* x,
                           int kernel, double beta0, double beta_br,
                           double beta_sh, double log_x0, double inv_log_r,
                           int bracket_mode, double* rhs, double* abs_sum,
                           double* m1_defect);

/* The same with the rows split over `threads` OpenMP threads (equal pair
 * counts), the per-thread double-double bins merged in thread order. */
int oracle_aggregation_rhs_threads(int n, const double* N, const double* x,
                                   int kernel, double beta0, double beta_br,
                                   double beta_sh, double log_x0, double inv_log_r,
                                   int bracket_mode, double* rhs, double* abs_sum,
                                   double* m1_defect, int threads);

int oracle_breakage_rhs(int n, int n_quad, const double* N, const double* x,
                        const double* t_q, const double* bw_q, int selection,
                        double S0, double v_ref, double alpha, double v_min,
                        double* rhs, double* abs_sum, double* m1_defect);

typedef struct {
    int n;
    const double* x;
    int use_agg;
    int kernel;
    double beta0, beta_br, beta_sh, log_x0, inv_log_r;
    int bracket_mode;
    int use_brk;
    int n_quad;
    const double* t_q;
    const double* bw_q;
    int selection;
    double S0, v_ref, alpha, v_min;
    int threads;          /* OpenMP threads of the aggregation sums (<= 1: serial) */
} oracle_problem_t;

/* classic RK4, `steps` steps of size dt, in place on Nst */
int oracle_rk4(const oracle_problem_t* p, double* Nst, double dt, int steps);

/* Plain-FP64 OpenMP restatement (pbe_cpu_port.c) used only as a timing
 * baseline: upper-triangle pair loop with thread-private rhs and a final
 * reduction, reference bracket rule. */
int port_aggregation_rhs(int n, const double* N, const double* x, int kernel,
                         double beta0, double beta_br, double beta_sh,
                         double log_x0, double inv_log_r, double* rhs,
                         int threads);
int port_breakage_rhs(int n, int n_quad, const double* N, const double* x,
                      const double* t_q, const double* bw_q, int selection,
                      double S0, double v_ref, double alpha, double v_min,
                      double* rhs, int threads);

#ifdef __cplusplus
}
#endif
#endif
