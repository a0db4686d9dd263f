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

/* Launch options of the aggregation kernels (no counterpart in the reference,
 * whose only knob is the compile-time PBE_TILE, src/aggregation_kernels.cuh:
 * 30-33).  tile_u = columns per lane: column blocks and diagonal-band tiles
 * are 64*tile_u wide; band_split = row slices per diagonal-band tile (each
 * slice one workgroup; offset segments per block for the 256-wide band
 * sweeps); form = 1: the band and the column strips as two launches (one
 * launch of every tile when the strips would hold too little work), 2: one
 * launch running both side by side (taken for tile_u = 4 up to 64 column
 * blocks, otherwise the same as 1), 3: one launch of every tile (the
 * diagonal tiles within the near-pair reach as row slices), 4: as 2 with
 * three workgroups resident per CU instead of two.  0 picks the library's
 * autotuned choice
 * for (n, kernel_type, grid reach), where the reach of a geometric grid
 * (inv_log_r > 0) is log_r(1/(r-1)) bins: how far below the diagonal a pair
 * can still land beyond its bracket (k, k+1). */
typedef struct pbe_agg_options {
    int tile_u;       /* 0, 1, 2, 4 or 8 */
    int band_split;   /* 0, 1, 2, 4, 8 or 16 */
    int form;         /* 0, 1, 2, 3 or 4 */
} pbe_agg_options;

/* pbe_launch_aggregation_rhs with explicit launch options (same results). */
hipError_t pbe_launch_aggregation_rhs_ex(const double* N, const double* x,
                                         double* rhs,
                                         const pbe_aggregation_params* params,
                                         const pbe_agg_options* options,
                                         hipStream_t stream);

/* The autotuned options for params (n, kernel_type and the grid's reach):
 * what tile_u, band_split and form = 0 resolve to.  hipErrorInvalidValue for
 * NULL pointers, n <= 0 or an unknown kernel_type. */
hipError_t pbe_aggregation_default_options(const pbe_aggregation_params* params,
                                           pbe_agg_options* options);

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

/* ------------------------------------------------------------------------ */
/* Atomic-free gather mode (no counterpart in the reference, whose FP64      */
/* atomics make the summation order, and so the last bits, vary from run to */
/* run).  Same contract and errors as the launches above, plus a caller-    */
/* owned device workspace of at least the returned size                     */
/* (hipErrorInvalidValue if smaller, hipErrorInvalidDevicePointer if NULL). */
/* Every rhs entry is written by one thread and every sum is formed in an   */
/* order fixed by n (and n_quad) alone: two calls on the same inputs give   */
/* identical bits.  Results agree with the atomic launches to rounding.     */
/* ------------------------------------------------------------------------ */
size_t pbe_aggregation_deterministic_workspace_bytes(int n);
size_t pbe_breakage_deterministic_workspace_bytes(int n);

hipError_t pbe_launch_aggregation_rhs_deterministic(const double* N, const double* x,
                                                    double* rhs,
                                                    const pbe_aggregation_params* params,
                                                    void* workspace, size_t workspace_bytes,
                                                    hipStream_t stream);

hipError_t pbe_launch_breakage_rhs_deterministic(const double* N, const double* x,
                                                 const double* t_q, const double* bw_q,
                                                 double* rhs,
                                                 const pbe_breakage_params* params,
                                                 void* workspace, size_t workspace_bytes,
                                                 hipStream_t stream);

/* pbe_launch_aggregation_rhs_deterministic without a workspace argument and
 * without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The workspace is a
 * static device array inside the library sized for n up to
 * PBE_DETERMINISTIC_NOALLOC_MAX_N (about 23 MB per device, held until the
 * process ends); n above that returns hipErrorInvalidValue before any launch.
 * All calls on one device share that array: calls that overlap on one device
 * (different streams or threads) race and must be serialised by the caller.
 * Same results, bit for bit, as the workspace form; other errors as there. */
#define PBE_DETERMINISTIC_NOALLOC_MAX_N 65536
hipError_t pbe_launch_aggregation_rhs_deterministic_noalloc(const double* N, const double* x,
                                                            double* rhs,
                                                            const pbe_aggregation_params* params,
                                                            hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Device time integration (no counterpart in the reference library: it     */
/* replaces the host-side euler_step / rk4_step of the reference's example  */
/* harness, examples/common/pbe_examples_utils.hpp:391-457, with the same   */
/* arithmetic and evaluation order, and no host round-trips).               */
/* ------------------------------------------------------------------------ */
enum { PBE_METHOD_RK4 = 0, PBE_METHOD_EULER = 1 };

/* Bytes of device workspace an integrator needs for n bins (RK4: 3n doubles,
 * Euler: n doubles); 0 for n <= 0. */
size_t pbe_integrator_workspace_bytes(int n, int method);

/*
 * Advance N in place by `steps` steps of size dt of dN/dt = R_agg + R_brk,
 * where agg or brk may be NULL (that source term is then absent).  x, t_q,
 * bw_q as for the launch functions; workspace: caller-owned device memory of
 * at least pbe_integrator_workspace_bytes(n, method) bytes.  Enqueue-only.
 */
hipError_t pbe_integrate(double* N, const double* x, const double* t_q,
                         const double* bw_q, const pbe_aggregation_params* agg,
                         const pbe_breakage_params* brk, double dt, int steps,
                         int method, void* workspace, size_t workspace_bytes,
                         hipStream_t stream);

/* A hipGraph holding `steps_per_launch` integrator steps (all pointers and
 * parameters are baked in at creation; the graph contains no memset node). */
typedef struct pbe_graph_s* pbe_graph;
hipError_t pbe_integrator_graph_create(pbe_graph* out, double* N, const double* x,
                                       const double* t_q, const double* bw_q,
                                       const pbe_aggregation_params* agg,
                                       const pbe_breakage_params* brk, double dt,
                                       int steps_per_launch, int method,
                                       void* workspace, size_t workspace_bytes);
hipError_t pbe_graph_launch(pbe_graph graph, hipStream_t stream);
hipError_t pbe_graph_destroy(pbe_graph graph);

/* ------------------------------------------------------------------------ */
/* Batched form: many independent instances of uneven n in one persistent   */
/* launch (no counterpart in the reference, which evaluates one instance    */
/* per call).  All instances share the aggregation kernel type, the         */
/* selection function and the daughter table (t_q, bw_q); each has its own  */
/* grid, coefficients and slice [offset, offset + n) of the packed N, x and */
/* rhs arrays.                                                              */
/* ------------------------------------------------------------------------ */
typedef struct pbe_batch_instance {
    int    n;          /* bins of this instance */
    int    offset;     /* index of its bin 0 in the packed arrays */
    double log_x0;     /* as AggregationParams */
    double inv_log_r;  /* 0: binary-search bracket */
    double beta0, beta_br, beta_sh;   /* aggregation coefficients */
    double S0, v_ref, alpha, v_min;   /* breakage selection parameters */
} pbe_batch_instance;

typedef struct pbe_batch_plan_s* pbe_batch_plan;

/* Build the work list (pair tiles and parent blocks of every instance,
 * longest first) and upload it with the descriptors.  kernel_type = -1 means
 * no aggregation, selection = -1 no breakage.  The plan owns device memory
 * until pbe_batch_plan_destroy. */
hipError_t pbe_batch_plan_create(pbe_batch_plan* out,
                                 const pbe_batch_instance* instances, int count,
                                 int kernel_type, int selection, int n_quad);
/* pbe_batch_plan_create with every instance's pair tiles 64*tile_u wide
 * (tile_u = 1, 2, 4 or 8; 0 picks the autotuned width per instance, as
 * pbe_batch_plan_create does).  Same results. */
hipError_t pbe_batch_plan_create_ex(pbe_batch_plan* out,
                                    const pbe_batch_instance* instances, int count,
                                    int kernel_type, int selection, int n_quad,
                                    int tile_u);
hipError_t pbe_batch_plan_destroy(pbe_batch_plan plan);
/* Length of the packed arrays the plan addresses (max offset + n). */
int pbe_batch_plan_size(pbe_batch_plan plan);

/* rhs += R_agg + R_brk of every instance (packed arrays; rhs zeroed by the
 * caller).  Enqueue-only. */
hipError_t pbe_launch_batched_rhs(pbe_batch_plan plan, const double* N,
                                  const double* x, const double* t_q,
                                  const double* bw_q, double* rhs,
                                  hipStream_t stream);

/* Integrate every instance of the plan (workspace sized for
 * pbe_batch_plan_size(plan) bins), eagerly or as a hipGraph. */
hipError_t pbe_batch_integrate(pbe_batch_plan plan, double* N, const double* x,
                               const double* t_q, const double* bw_q, double dt,
                               int steps, int method, void* workspace,
                               size_t workspace_bytes, hipStream_t stream);
hipError_t pbe_batch_integrator_graph_create(pbe_graph* out, pbe_batch_plan plan,
                                             double* N, const double* x,
                                             const double* t_q, const double* bw_q,
                                             double dt, int steps_per_launch,
                                             int method, void* workspace,
                                             size_t workspace_bytes);

/* ------------------------------------------------------------------------ */
/* Differentiation (no counterpart in the reference): hand-written          */
/* vector-Jacobian products of the right-hand side and the discrete adjoint */
/* of explicit RK4 with checkpoint-and-recompute.  Gradient slots of        */
/* `gtheta` (5 doubles, device memory): d/d beta0, beta_br, beta_sh, S0,    */
/* alpha.  All results are atomic-free and bitwise repeatable.              */
/* ------------------------------------------------------------------------ */
enum { PBE_GRAD_BETA0 = 0, PBE_GRAD_BETA_BR = 1, PBE_GRAD_BETA_SH = 2,
       PBE_GRAD_S0 = 3, PBE_GRAD_ALPHA = 4, PBE_GRAD_COUNT = 5 };

size_t pbe_vjp_workspace_bytes(int n);

/* gN = lam^T dR/dN and gtheta = lam^T dR/dtheta at N (both overwritten),
 * R = R_agg + R_brk as for pbe_integrate. */
hipError_t pbe_rhs_vjp(const double* N, const double* x, const double* t_q,
                       const double* bw_q, const double* lam,
                       const pbe_aggregation_params* agg,
                       const pbe_breakage_params* brk, double* gN, double* gtheta,
                       void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Workspace of the RK4 adjoint: ceil(steps/every) checkpoints plus the 4
 * stage states of one segment of `every` steps, plus O(n) scratch.
 * checkpoint_every = PBE_STORE_ALL_STAGES (0) keeps the 4 stage states of
 * every step instead (4 steps n doubles): the reverse pass then recomputes
 * nothing. */
enum { PBE_STORE_ALL_STAGES = 0 };
size_t pbe_adjoint_workspace_bytes(int n, int steps, int checkpoint_every);

/* Forward pass: advance N in place by `steps` RK4 steps, storing a
 * checkpoint every `checkpoint_every` steps (or every stage state, for
 * PBE_STORE_ALL_STAGES) in the workspace. */
hipError_t pbe_adjoint_rk4_forward(double* N, const double* x, const double* t_q,
                                   const double* bw_q,
                                   const pbe_aggregation_params* agg,
                                   const pbe_breakage_params* brk, double dt,
                                   int steps, int checkpoint_every, void* workspace,
                                   size_t workspace_bytes, hipStream_t stream);

/* Reverse pass after pbe_adjoint_rk4_forward with the same problem and
 * workspace: given gN_final = dL/dN(t_end), writes gN0 = dL/dN(0) and
 * gtheta = dL/dtheta, recomputing each segment from its checkpoint. */
hipError_t pbe_adjoint_rk4_backward(const double* gN_final, double* gN0,
                                    double* gtheta, const double* x,
                                    const double* t_q, const double* bw_q,
                                    const pbe_aggregation_params* agg,
                                    const pbe_breakage_params* brk, double dt,
                                    int steps, int checkpoint_every,
                                    void* workspace, size_t workspace_bytes,
                                    hipStream_t stream);

/* Batched VJP and RK4 adjoint over a plan (workspace sized for
 * pbe_batch_plan_size(plan) bins).  The instances' coefficients are fitted
 * jointly: gtheta slots 0-3 hold sum over instances of
 * coef_i dL/dcoef_i = dL/d ln(c) for a common scale c of beta0, beta_br,
 * beta_sh and S0; slot 4 holds dL/dalpha. */
hipError_t pbe_batch_rhs_vjp(pbe_batch_plan plan, const double* N, const double* x,
                             const double* t_q, const double* bw_q, const double* lam,
                             double* gN, double* gtheta, void* workspace,
                             size_t workspace_bytes, hipStream_t stream);
hipError_t pbe_batch_adjoint_rk4_forward(pbe_batch_plan plan, double* N, const double* x,
                                         const double* t_q, const double* bw_q, double dt,
                                         int steps, int checkpoint_every, void* workspace,
                                         size_t workspace_bytes, hipStream_t stream);
hipError_t pbe_batch_adjoint_rk4_backward(pbe_batch_plan plan, const double* gN_final,
                                          double* gN0, double* gtheta, const double* x,
                                          const double* t_q, const double* bw_q, double dt,
                                          int steps, int checkpoint_every, void* workspace,
                                          size_t workspace_bytes, hipStream_t stream);

/* Library version string, e.g. "pbe_hip 0.1.0 gfx950". */
const char* pbe_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PBE_HIP_H */
