/*
 * pbe_hip.h -- C ABI of the MI355X (gfx950) fixed-pivot PBE source terms.
 *
 * Drop-in boundary for acuoci/pbe-fixed-pivot-cuda: the two reference entry
 * points become two extern "C" functions over plain pointers, POD parameter
 * structs with the reference's exact field order, and hipStream_t.  The C++
 * names of the reference (namespace pbe_cuda, AggregationParams, ...) are
 * provided as inline forwarders in <pbe_cuda/pbe_cuda.cuh>.
 *
 * Contract kept from the reference (include/pbe_cuda/aggregation.cuh:79-102,
 * include/pbe_cuda/breakage.cuh:81-121):
 *   - all pointers are caller-owned device memory, the library allocates
 *     nothing and keeps no state between calls;
 *   - rhs is accumulated into (rhs += source); the caller zeroes it once and
 *     may call aggregation and breakage in either order;
 *   - the call only enqueues work on `stream` and never synchronises; the
 *     return value is the launch status;
 *   - argument checks happen before any HIP runtime call, in the reference's
 *     order: n, then pointers, then the enum.
 */
#ifndef PBE_HIP_H
#define PBE_HIP_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AggregationKernel of include/pbe_cuda/aggregation.cuh:32-39 */
enum {
    PBE_AGG_CONSTANT       = 0,
    PBE_AGG_SUM            = 1,
    PBE_AGG_PRODUCT        = 2,
    PBE_AGG_BROWNIAN       = 3,
    PBE_AGG_SHEAR          = 4,
    PBE_AGG_BROWNIAN_SHEAR = 5
};

/* BreakageSelection of include/pbe_cuda/breakage.cuh:30-35 */
enum {
    PBE_SEL_CONSTANT  = 0,
    PBE_SEL_LINEAR    = 1,
    PBE_SEL_POWER_LAW = 2,
    PBE_SEL_THRESHOLD = 3
};

/* Layout-identical mirror of AggregationParams
 * (include/pbe_cuda/aggregation.cuh:59-73): sizeof 64,
 * offsets 0, 8, 16, 24, 32, 40, 48, 56. */
typedef struct pbe_aggregation_params {
    int    n;
    double log_x0;
    double inv_log_r;     /* 0 selects the binary-search bracket path */
    int    kernel_type;   /* PBE_AGG_* */
    double beta0;
    double beta_br;
    double beta_sh;
    int    block_size;    /* tuning hint; <= 0 means the default */
} pbe_aggregation_params;

/* Layout-identical mirror of BreakageParams
 * (include/pbe_cuda/breakage.cuh:63-77): sizeof 56,
 * offsets 0, 4, 8, 16, 24, 32, 40, 48. */
typedef struct pbe_breakage_params {
    int    n;
    int    n_quad;
    int    selection;     /* PBE_SEL_* */
    double S0;
    double v_ref;
    double alpha;
    double v_min;
    int    block_size;    /* tuning hint; must be in [1, 1024] */
} pbe_breakage_params;

/*
 * rhs[0..n) += fixed-pivot aggregation source of N on pivots x.
 * Replaces pbe_cuda::launch_aggregation_rhs
 * (include/pbe_cuda/aggregation.cuh:115-120, src/aggregation.cu:16-61).
 * Errors: hipErrorInvalidValue for params == NULL, n <= 0 or an unknown
 * kernel_type; hipErrorInvalidDevicePointer for a NULL array;
 * hipErrorInvalidConfiguration for block_size > 1024 (the reference's launch
 * fails the same way); otherwise hipGetLastError() after the launch.
 * Non-deterministic summation order (FP64 atomics).
 */
hipError_t pbe_launch_aggregation_rhs(const double* N, const double* x,
                                      double* rhs,
                                      const pbe_aggregation_params* params,
                                      hipStream_t stream);

/*
 * rhs[0..n) += fixed-pivot breakage source (death and quadrature birth).
 * Replaces pbe_cuda::launch_breakage_rhs
 * (include/pbe_cuda/breakage.cuh:122-129, src/breakage.cu:14-78).
 * Errors: hipErrorInvalidValue for params == NULL, n <= 0, n_quad <= 0, an
 * unknown selection or block_size <= 0 (the reference divides by zero there);
 * hipErrorInvalidDevicePointer for a NULL array; hipErrorInvalidConfiguration
 * for block_size > 1024.
 */
hipError_t pbe_launch_breakage_rhs(const double* N, const double* x,
                                   const double* t_q, const double* bw_q,
                                   double* rhs,
                                   const pbe_breakage_params* params,
                                   hipStream_t stream);

/* Library version string, e.g. "pbe_hip 0.1.0 gfx950". */
const char* pbe_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PBE_HIP_H */
