// pbe_internal.hpp -- launchers shared between the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "pbe_hip/pbe_hip.h"

namespace pbe {

// Enqueue the aggregation tiles (U = columns per lane: 2, 4 or 8).
// form 2: band and strips in one launch where the shape allows it (tile_u
// = 4, at most 64 column blocks); 1: two launches; 3: one launch of every
// tile; 4: as 2, three workgroups per CU.
hipError_t launch_aggregation(const double* N, const double* x, double* rhs,
                              const pbe_aggregation_params& p, int tile_u, int band_split,
                              int form, hipStream_t stream);

// Enqueue the fused breakage death + birth kernel.
hipError_t launch_breakage(const double* N, const double* x, const double* t_q,
                           const double* bw_q, double* rhs,
                           const pbe_breakage_params& p, hipStream_t stream);

// Band split the launchers use for a requested one (1, 2, 4, 8, 16; else 4).
inline int band_split_or_default(int s)
{
    return (s == 1 || s == 2 || s == 8 || s == 16) ? s : 4;
}

// The single tile launch (agg_tiles_kernel) an aggregation call at this
// shape makes: its work items and the leading tile diagonals run as row
// slices; items = -1 when the call makes other launches (strips, the fused
// band + strip launch).
struct TilesShape { long long items; int diags; };
TilesShape aggregation_tiles_shape(const pbe_aggregation_params& p, int tile_u, int band_split,
                                   int form);

// R_agg(N) + R_brk(N) added into rhs in one launch (rhs_fused.hip), for
// problems with both terms whose aggregation is a single tile launch.
// Returns hipErrorNotSupported, enqueueing nothing, for any other problem
// (the caller then makes the separate launches).
struct RhsProblem;
hipError_t launch_rhs_fused(const RhsProblem& p, const double* N, double* rhs,
                            hipStream_t stream);

// Launch shape of the aggregation kernels for a problem (the autotuned
// table, capi.hip): columns per lane, band-tile row slices and launch form.
struct TileChoice { int u, w, f; };
TileChoice tile_choice(int n, int kernel_type, double inv_log_r);
// Far-pair reach log_r(1 / (r - 1)) in bins of a geometric grid (>= 0), -1
// for inv_log_r = 0 (no geometric lookup).
double grid_reach(double inv_log_r);

// Atomic-free, bitwise repeatable forms (det_kernels.hip); ws holds
// det_*_workspace_doubles(n) doubles.
size_t det_aggregation_workspace_doubles(int n);
size_t det_breakage_workspace_doubles(int n);
hipError_t launch_aggregation_deterministic(const double* N, const double* x, double* rhs,
                                            const pbe_aggregation_params& p, double* ws,
                                            hipStream_t stream);
// The aggregation gather on the library's static per-device workspace (no
// allocation); n <= kDetNoAllocMaxN.
constexpr int kDetNoAllocMaxN = PBE_DETERMINISTIC_NOALLOC_MAX_N;
hipError_t launch_aggregation_deterministic_noalloc(const double* N, const double* x,
                                                    double* rhs,
                                                    const pbe_aggregation_params& p,
                                                    hipStream_t stream);
hipError_t launch_breakage_deterministic(const double* N, const double* x, const double* t_q,
                                         const double* bw_q, double* rhs,
                                         const pbe_breakage_params& p, double* ws,
                                         hipStream_t stream);

// One right-hand side R(N) = R_agg(N) + R_brk(N); either part may be absent.
// With `batch` set, R is the batched source of every instance of the plan and
// n is the packed length.
struct RhsProblem {
    int n;
    const double* x;
    const double* t_q;
    const double* bw_q;
    const pbe_aggregation_params* agg;
    const pbe_breakage_params* brk;
    pbe_batch_plan batch;
};

hipError_t batch_plan_create(pbe_batch_plan* out, const pbe_batch_instance* instances,
                             int count, int kernel_type, int selection, int n_quad,
                             int tile_u);
// Columns per lane of one batched instance's pair tiles (capi.hip).
int batch_tile_u(int n, int kernel_type, double inv_log_r);
hipError_t batch_plan_destroy(pbe_batch_plan plan);
int batch_plan_total(pbe_batch_plan plan);
bool batch_plan_needs_table(pbe_batch_plan plan);
hipError_t enqueue_batched_rhs(pbe_batch_plan plan, const double* N, const double* x,
                               const double* t_q, const double* bw_q, double* rhs,
                               hipStream_t s);

hipError_t enqueue_fill_zero(double* a, int n, hipStream_t s);
// rhs (zero on entry) += R(Nin)
hipError_t enqueue_rhs(const RhsProblem& p, const double* Nin, double* rhs, hipStream_t s);
// One classic RK4 step in place on N; work = 3n doubles whose first n (k) are
// zero on entry and are left zero on exit.
hipError_t enqueue_rk4_step(const RhsProblem& p, double* N, double dt, double* work,
                            hipStream_t s);
// As above, with the stage states Y2, Y3, Y4 written to stages[0..2] (NULL:
// a scratch slot of work).
hipError_t enqueue_rk4_step_stages(const RhsProblem& p, double* N, double dt, double* work,
                                   double* const* stages, hipStream_t s);
// As above from Nin to Nout (Nin is left as it is unless it is Nout), and
// Nin also copied to ysave (when non-null) by the first stage kernel.
hipError_t enqueue_rk4_step_io(const RhsProblem& p, const double* Nin, double* Nout, double dt,
                               double* work, double* const* stages, double* ysave,
                               hipStream_t s);
// One explicit Euler step; work = n doubles, zero on entry and exit.
hipError_t enqueue_euler_step(const RhsProblem& p, double* N, double dt, double* work,
                              hipStream_t s);

// ----------------------------------------------------------- adjoint ----
hipError_t enqueue_copy(const double* src, double* dst, int n, hipStream_t s);
hipError_t enqueue_vjp_table(const double* x, double* tab, int n, hipStream_t s);
// gN = lam . dR/dN at N (overwritten); gth_rows[n][5] += per-row parameter
// partials (beta0, beta_br, beta_sh, S0, alpha); tab from enqueue_vjp_table.
// part: scratch of vjp_part_doubles(n, batched) doubles.
// fuse (optional) folds one RK4 adjoint stage into the VJP's last pass:
//   lam_acc += gN;  av_next = c1 lam_base + c2 gN  (av_next may be null).
struct AdjointStageFuse {
    double* lam_acc;
    double* av_next;
    const double* lam_base;
    double c1, c2;
    double* lam_copy = nullptr;   // last stage: start of the next step
    double* av_start = nullptr;
    double c_start = 0.0;
};
hipError_t enqueue_vjp(const RhsProblem& p, const double* N, const double* lam,
                       const double* tab, double* gN, double* gth_rows, double* part,
                       hipStream_t s, const AdjointStageFuse* fuse = nullptr);
size_t vjp_part_doubles(int n, bool batched);

// Workspace of the checkpointed RK4 adjoint, offsets in doubles.
struct AdjointLayout {
    bool store_all;
    int nseg;
    // seg2, rk2, mu2: the second segment buffer and the integrator scratch of
    // the recompute running beside the reverse sweep (checkpoint mode only)
    size_t ckpt, seg, seg2, tab, lam, lam2, av, mu, rk, rk2, mu2, gth, part, total;
};
AdjointLayout adjoint_layout(int n, int steps, int every);
hipError_t enqueue_adjoint_forward(const RhsProblem& p, double* N, double dt, int steps,
                                   int every, double* ws, hipStream_t s);
hipError_t enqueue_adjoint_backward(const RhsProblem& p, const double* gN_final, double* gN0,
                                    double* gtheta, double dt, int steps, int every,
                                    double* ws, hipStream_t s);
hipError_t enqueue_gtheta_reduce(const double* rows, double* out, int n, hipStream_t s);

}  // namespace pbe
