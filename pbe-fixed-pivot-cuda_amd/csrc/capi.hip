// capi.hip -- extern "C" boundary: argument checks, dispatch, launch status.
//
// Check order and error codes follow src/aggregation.cu:23-24,54 and
// src/breakage.cu:23-24,46,74 of the reference (n, then pointers, then the
// enum).  Every check runs before the first HIP runtime call.
#include <algorithm>
#include <cmath>

#include "pbe_internal.hpp"

static_assert(sizeof(pbe_aggregation_params) == 64, "AggregationParams layout");
static_assert(offsetof(pbe_aggregation_params, log_x0) == 8, "layout");
static_assert(offsetof(pbe_aggregation_params, inv_log_r) == 16, "layout");
static_assert(offsetof(pbe_aggregation_params, kernel_type) == 24, "layout");
static_assert(offsetof(pbe_aggregation_params, beta0) == 32, "layout");
static_assert(offsetof(pbe_aggregation_params, beta_br) == 40, "layout");
static_assert(offsetof(pbe_aggregation_params, beta_sh) == 48, "layout");
static_assert(offsetof(pbe_aggregation_params, block_size) == 56, "layout");
static_assert(sizeof(pbe_breakage_params) == 56, "BreakageParams layout");
static_assert(offsetof(pbe_breakage_params, n_quad) == 4, "layout");
static_assert(offsetof(pbe_breakage_params, selection) == 8, "layout");
static_assert(offsetof(pbe_breakage_params, S0) == 16, "layout");
static_assert(offsetof(pbe_breakage_params, v_ref) == 24, "layout");
static_assert(offsetof(pbe_breakage_params, alpha) == 32, "layout");
static_assert(offsetof(pbe_breakage_params, v_min) == 40, "layout");
static_assert(offsetof(pbe_breakage_params, block_size) == 48, "layout");

namespace pbe {

// Columns per lane and band-tile row slices of the aggregation kernels, from
// the autotuned table (tile_table.inc: rows {n_max, kernel_type, reach_max,
// tile_u, band_split, form}, written by pbe-fixed-pivot-cuda_amd/autotune.py from
// timings on an MI355X).  The reach of a geometric grid, log_r(1 / (r - 1))
// bins, is how far below the diagonal a pair (j, k) can still be near (its
// birth beyond the bracket (k, k+1)); reach_max -1 rows are for grids without
// the geometric lookup (inv_log_r = 0).  First match wins.
struct TileRule { int n_max, kernel, reach_max, u, w, f; };
static const TileRule kTileTable[] = {
#include "tile_table.inc"
};

double grid_reach(double inv_log_r)
{
    if (!(inv_log_r > 0.0) || !std::isfinite(inv_log_r)) return -1.0;
    const double rm1 = std::expm1(1.0 / inv_log_r);            // r - 1
    if (!(rm1 > 0.0)) return -1.0;
    const double reach = -std::log(rm1) * inv_log_r;
    return reach > 0.0 ? reach : 0.0;
}

// Tile width of the batched form's pair tiles (batch_tile_table.inc: rows
// {n_max, kernel_type, reach_max, tile_u}, from the same tuner; the band
// tiles always run as four row slices there).
struct BatchTileRule { int n_max, kernel, reach_max, u; };
static const BatchTileRule kBatchTileTable[] = {
#include "batch_tile_table.inc"
};

static bool reach_matches(double reach, int reach_max)
{
    return reach < 0.0 ? reach_max < 0 : (reach_max >= 0 && reach <= reach_max);
}

int batch_tile_u(int n, int kernel_type, double inv_log_r)
{
    const double reach = grid_reach(inv_log_r);
    for (const BatchTileRule& r : kBatchTileTable)
        if (n <= r.n_max && (r.kernel < 0 || r.kernel == kernel_type) &&
            reach_matches(reach, r.reach_max))
            return r.u;
    return 2;
}

TileChoice tile_choice(int n, int kernel_type, double inv_log_r)
{
    const double reach = grid_reach(inv_log_r);
    for (const TileRule& r : kTileTable)
        if (n <= r.n_max && (r.kernel < 0 || r.kernel == kernel_type) &&
            reach_matches(reach, r.reach_max))
            return TileChoice{r.u, r.w, r.f};
    return TileChoice{4, 4, 1};
}

}  // namespace pbe

extern "C" hipError_t pbe_launch_aggregation_rhs(const double* N, const double* x,
                                                 double* rhs,
                                                 const pbe_aggregation_params* p,
                                                 hipStream_t stream)
{
    if (!p || p->n <= 0) return hipErrorInvalidValue;
    if (!N || !x || !rhs) return hipErrorInvalidDevicePointer;
    if (p->kernel_type < 0 || p->kernel_type > 5) return hipErrorInvalidValue;
    if (p->block_size > 1024) return hipErrorInvalidConfiguration;
    const pbe::TileChoice t = pbe::tile_choice(p->n, p->kernel_type, p->inv_log_r);
    hipError_t e = pbe::launch_aggregation(N, x, rhs, *p, t.u, t.w, t.f, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" hipError_t pbe_launch_aggregation_rhs_ex(const double* N, const double* x,
                                                    double* rhs,
                                                    const pbe_aggregation_params* p,
                                                    const pbe_agg_options* opts,
                                                    hipStream_t stream)
{
    if (!p || p->n <= 0) return hipErrorInvalidValue;
    if (!N || !x || !rhs) return hipErrorInvalidDevicePointer;
    if (p->kernel_type < 0 || p->kernel_type > 5) return hipErrorInvalidValue;
    if (p->block_size > 1024) return hipErrorInvalidConfiguration;
    const pbe::TileChoice t = pbe::tile_choice(p->n, p->kernel_type, p->inv_log_r);
    int u = opts ? opts->tile_u : 0;
    int w = opts ? opts->band_split : 0;
    int f = opts ? opts->form : 0;
    if (u == 0) u = t.u;
    if (w == 0) w = t.w;
    if (f == 0) f = t.f;
    if (u != 1 && u != 2 && u != 4 && u != 8) return hipErrorInvalidValue;
    if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return hipErrorInvalidValue;
    if (f < 1 || f > 4) return hipErrorInvalidValue;
    hipError_t e = pbe::launch_aggregation(N, x, rhs, *p, u, w, f, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" hipError_t pbe_aggregation_default_options(const pbe_aggregation_params* p,
                                                      pbe_agg_options* out)
{
    if (!p || !out || p->n <= 0) return hipErrorInvalidValue;
    if (p->kernel_type < 0 || p->kernel_type > 5) return hipErrorInvalidValue;
    const pbe::TileChoice t = pbe::tile_choice(p->n, p->kernel_type, p->inv_log_r);
    out->tile_u = t.u;
    out->band_split = t.w;
    out->form = t.f;
    return hipSuccess;
}

extern "C" hipError_t pbe_launch_breakage_rhs(const double* N, const double* x,
                                              const double* t_q, const double* bw_q,
                                              double* rhs,
                                              const pbe_breakage_params* p,
                                              hipStream_t stream)
{
    if (!p || p->n <= 0 || p->n_quad <= 0) return hipErrorInvalidValue;
    if (!N || !x || !t_q || !bw_q || !rhs) return hipErrorInvalidDevicePointer;
    if (p->selection < 0 || p->selection > 3) return hipErrorInvalidValue;
    if (p->block_size <= 0) return hipErrorInvalidValue;
    if (p->block_size > 1024) return hipErrorInvalidConfiguration;
    hipError_t e = pbe::launch_breakage(N, x, t_q, bw_q, rhs, *p, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// ------------------------------------------------- deterministic gather ----
extern "C" size_t pbe_aggregation_deterministic_workspace_bytes(int n)
{
    if (n <= 0) return 0;
    return pbe::det_aggregation_workspace_doubles(n) * sizeof(double);
}

extern "C" size_t pbe_breakage_deterministic_workspace_bytes(int n)
{
    if (n <= 0) return 0;
    return pbe::det_breakage_workspace_doubles(n) * sizeof(double);
}

extern "C" hipError_t pbe_launch_aggregation_rhs_deterministic(
    const double* N, const double* x, double* rhs, const pbe_aggregation_params* p,
    void* workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (!p || p->n <= 0) return hipErrorInvalidValue;
    if (!N || !x || !rhs || !workspace) return hipErrorInvalidDevicePointer;
    if (p->kernel_type < 0 || p->kernel_type > 5) return hipErrorInvalidValue;
    if (p->block_size > 1024) return hipErrorInvalidConfiguration;
    if (workspace_bytes < pbe_aggregation_deterministic_workspace_bytes(p->n))
        return hipErrorInvalidValue;
    hipError_t e = pbe::launch_aggregation_deterministic(N, x, rhs, *p,
                                                         static_cast<double*>(workspace), stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" hipError_t pbe_launch_aggregation_rhs_deterministic_noalloc(
    const double* N, const double* x, double* rhs, const pbe_aggregation_params* p,
    hipStream_t stream)
{
    if (!p || p->n <= 0) return hipErrorInvalidValue;
    if (!N || !x || !rhs) return hipErrorInvalidDevicePointer;
    if (p->kernel_type < 0 || p->kernel_type > 5) return hipErrorInvalidValue;
    if (p->block_size > 1024) return hipErrorInvalidConfiguration;
    if (p->n > PBE_DETERMINISTIC_NOALLOC_MAX_N) return hipErrorInvalidValue;
    hipError_t e = pbe::launch_aggregation_deterministic_noalloc(N, x, rhs, *p, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" hipError_t pbe_launch_breakage_rhs_deterministic(
    const double* N, const double* x, const double* t_q, const double* bw_q, double* rhs,
    const pbe_breakage_params* p, void* workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (!p || p->n <= 0 || p->n_quad <= 0) return hipErrorInvalidValue;
    if (!N || !x || !t_q || !bw_q || !rhs || !workspace) return hipErrorInvalidDevicePointer;
    if (p->selection < 0 || p->selection > 3) return hipErrorInvalidValue;
    if (p->block_size <= 0) return hipErrorInvalidValue;
    if (p->block_size > 1024) return hipErrorInvalidConfiguration;
    if (workspace_bytes < pbe_breakage_deterministic_workspace_bytes(p->n))
        return hipErrorInvalidValue;
    hipError_t e = pbe::launch_breakage_deterministic(N, x, t_q, bw_q, rhs, *p,
                                                      static_cast<double*>(workspace), stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" const char* pbe_version(void)
{
    return "pbe_hip 0.1.0 gfx950";
}

// ---------------------------------------------------------------------------
// Device time integration and its hipGraph (no counterpart in the reference
// library; replaces the host loop of examples/common/pbe_examples_utils.hpp:
// 391-457).
// ---------------------------------------------------------------------------
namespace {

hipError_t make_problem(pbe::RhsProblem& p, const double* x, const double* t_q,
                        const double* bw_q, const pbe_aggregation_params* agg,
                        const pbe_breakage_params* brk)
{
    if (!agg && !brk) return hipErrorInvalidValue;
    const int n = agg ? agg->n : brk->n;
    if (n <= 0) return hipErrorInvalidValue;
    if (agg && brk && agg->n != brk->n) return hipErrorInvalidValue;
    if (brk && brk->n_quad <= 0) return hipErrorInvalidValue;
    if (!x) return hipErrorInvalidDevicePointer;
    if (brk && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    if (agg && (agg->kernel_type < 0 || agg->kernel_type > 5)) return hipErrorInvalidValue;
    if (brk && (brk->selection < 0 || brk->selection > 3)) return hipErrorInvalidValue;
    p = pbe::RhsProblem{n, x, t_q, bw_q, agg, brk, nullptr};
    return hipSuccess;
}

}  // namespace

extern "C" size_t pbe_integrator_workspace_bytes(int n, int method)
{
    if (n <= 0) return 0;
    return (method == PBE_METHOD_EULER ? 1u : 3u) * (size_t)n * sizeof(double);
}

static hipError_t integrate_problem(const pbe::RhsProblem& p, double* N, double dt,
                                   int steps, int method, void* workspace,
                                   size_t workspace_bytes, hipStream_t stream)
{
    hipError_t e;
    if (method != PBE_METHOD_RK4 && method != PBE_METHOD_EULER) return hipErrorInvalidValue;
    if (steps < 0) return hipErrorInvalidValue;
    if (!N || !workspace) return hipErrorInvalidDevicePointer;
    if (workspace_bytes < pbe_integrator_workspace_bytes(p.n, method)) return hipErrorInvalidValue;
    double* w = static_cast<double*>(workspace);
    if ((e = pbe::enqueue_fill_zero(w, p.n, stream)) != hipSuccess) return e;
    for (int s = 0; s < steps; ++s) {
        e = method == PBE_METHOD_RK4 ? pbe::enqueue_rk4_step(p, N, dt, w, stream)
                                     : pbe::enqueue_euler_step(p, N, dt, w, stream);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

extern "C" hipError_t pbe_integrate(double* N, const double* x, const double* t_q,
                                    const double* bw_q, const pbe_aggregation_params* agg,
                                    const pbe_breakage_params* brk, double dt, int steps,
                                    int method, void* workspace, size_t workspace_bytes,
                                    hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = make_problem(p, x, t_q, bw_q, agg, brk);
    if (e != hipSuccess) return e;
    return integrate_problem(p, N, dt, steps, method, workspace, workspace_bytes, stream);
}

struct pbe_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static hipError_t graph_problem(pbe_graph* out, const pbe::RhsProblem& p, double* N,
                               double dt, int steps_per_launch, int method, void* workspace,
                               size_t workspace_bytes)
{
    hipError_t e;
    if (method != PBE_METHOD_RK4 && method != PBE_METHOD_EULER) return hipErrorInvalidValue;
    if (steps_per_launch <= 0) return hipErrorInvalidValue;
    if (!N || !workspace) return hipErrorInvalidDevicePointer;
    if (workspace_bytes < pbe_integrator_workspace_bytes(p.n, method)) return hipErrorInvalidValue;

    hipStream_t s = nullptr;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return e;
    auto* g = new pbe_graph_s();
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        double* w = static_cast<double*>(workspace);
        hipError_t ee = pbe::enqueue_fill_zero(w, p.n, s);
        for (int k = 0; ee == hipSuccess && k < steps_per_launch; ++k)
            ee = method == PBE_METHOD_RK4 ? pbe::enqueue_rk4_step(p, N, dt, w, s)
                                          : pbe::enqueue_euler_step(p, N, dt, w, s);
        e = hipStreamEndCapture(s, &g->graph);
        if (ee != hipSuccess) e = ee;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return e;
    }
    *out = g;
    return hipSuccess;
}

extern "C" hipError_t pbe_integrator_graph_create(pbe_graph* out, double* N, const double* x,
                                                  const double* t_q, const double* bw_q,
                                                  const pbe_aggregation_params* agg,
                                                  const pbe_breakage_params* brk, double dt,
                                                  int steps_per_launch, int method,
                                                  void* workspace, size_t workspace_bytes)
{
    if (!out) return hipErrorInvalidValue;
    *out = nullptr;
    pbe::RhsProblem p;
    hipError_t e = make_problem(p, x, t_q, bw_q, agg, brk);
    if (e != hipSuccess) return e;
    return graph_problem(out, p, N, dt, steps_per_launch, method, workspace, workspace_bytes);
}

// ------------------------------------------------------------- batched ----
extern "C" hipError_t pbe_batch_plan_create(pbe_batch_plan* out,
                                            const pbe_batch_instance* instances, int count,
                                            int kernel_type, int selection, int n_quad)
{
    return pbe::batch_plan_create(out, instances, count, kernel_type, selection, n_quad, 0);
}

extern "C" hipError_t pbe_batch_plan_create_ex(pbe_batch_plan* out,
                                               const pbe_batch_instance* instances, int count,
                                               int kernel_type, int selection, int n_quad,
                                               int tile_u)
{
    return pbe::batch_plan_create(out, instances, count, kernel_type, selection, n_quad, tile_u);
}

extern "C" hipError_t pbe_batch_plan_destroy(pbe_batch_plan plan)
{
    return pbe::batch_plan_destroy(plan);
}

extern "C" int pbe_batch_plan_size(pbe_batch_plan plan) { return pbe::batch_plan_total(plan); }

static hipError_t batch_problem(pbe::RhsProblem& p, pbe_batch_plan plan, const double* x,
                                const double* t_q, const double* bw_q)
{
    if (!plan) return hipErrorInvalidValue;
    if (!x) return hipErrorInvalidDevicePointer;
    p = pbe::RhsProblem{pbe::batch_plan_total(plan), x, t_q, bw_q, nullptr, nullptr, plan};
    return hipSuccess;
}

extern "C" hipError_t pbe_launch_batched_rhs(pbe_batch_plan plan, const double* N,
                                             const double* x, const double* t_q,
                                             const double* bw_q, double* rhs,
                                             hipStream_t stream)
{
    if (!plan) return hipErrorInvalidValue;
    if (!N || !x || !rhs) return hipErrorInvalidDevicePointer;
    if (pbe::batch_plan_needs_table(plan) && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    return pbe::enqueue_batched_rhs(plan, N, x, t_q, bw_q, rhs, stream);
}

extern "C" hipError_t pbe_batch_integrate(pbe_batch_plan plan, double* N, const double* x,
                                          const double* t_q, const double* bw_q, double dt,
                                          int steps, int method, void* workspace,
                                          size_t workspace_bytes, hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = batch_problem(p, plan, x, t_q, bw_q);
    if (e != hipSuccess) return e;
    if (pbe::batch_plan_needs_table(plan) && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    return integrate_problem(p, N, dt, steps, method, workspace, workspace_bytes, stream);
}

extern "C" hipError_t pbe_batch_integrator_graph_create(pbe_graph* out, pbe_batch_plan plan,
                                                        double* N, const double* x,
                                                        const double* t_q, const double* bw_q,
                                                        double dt, int steps_per_launch,
                                                        int method, void* workspace,
                                                        size_t workspace_bytes)
{
    if (!out) return hipErrorInvalidValue;
    *out = nullptr;
    pbe::RhsProblem p;
    hipError_t e = batch_problem(p, plan, x, t_q, bw_q);
    if (e != hipSuccess) return e;
    if (pbe::batch_plan_needs_table(plan) && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    return graph_problem(out, p, N, dt, steps_per_launch, method, workspace, workspace_bytes);
}

extern "C" hipError_t pbe_graph_launch(pbe_graph g, hipStream_t stream)
{
    if (!g || !g->exec) return hipErrorInvalidValue;
    return hipGraphLaunch(g->exec, stream);
}

extern "C" hipError_t pbe_graph_destroy(pbe_graph g)
{
    if (!g) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    if (g->exec) e = hipGraphExecDestroy(g->exec);
    if (g->graph) {
        hipError_t e2 = hipGraphDestroy(g->graph);
        if (e == hipSuccess) e = e2;
    }
    delete g;
    return e;
}

// --------------------------------------------------- VJP and RK4 adjoint ----
extern "C" size_t pbe_vjp_workspace_bytes(int n)
{
    if (n <= 0) return 0;
    const size_t part = std::max(pbe::vjp_part_doubles(n, false), pbe::vjp_part_doubles(n, true));
    return ((size_t)9 * (size_t)n + part) * sizeof(double);
}

extern "C" hipError_t pbe_rhs_vjp(const double* N, const double* x, const double* t_q,
                                  const double* bw_q, const double* lam,
                                  const pbe_aggregation_params* agg,
                                  const pbe_breakage_params* brk, double* gN, double* gtheta,
                                  void* workspace, size_t workspace_bytes, hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = make_problem(p, x, t_q, bw_q, agg, brk);
    if (e != hipSuccess) return e;
    if (!N || !lam || !gN || !gtheta || !workspace) return hipErrorInvalidDevicePointer;
    if (workspace_bytes < pbe_vjp_workspace_bytes(p.n)) return hipErrorInvalidValue;
    double* tab = static_cast<double*>(workspace);
    double* rows = tab + 4 * (size_t)p.n;
    double* part = rows + 5 * (size_t)p.n;
    if ((e = pbe::enqueue_vjp_table(x, tab, p.n, stream)) != hipSuccess) return e;
    if ((e = pbe::enqueue_fill_zero(rows, 5 * p.n, stream)) != hipSuccess) return e;
    if ((e = pbe::enqueue_vjp(p, N, lam, tab, gN, rows, part, stream)) != hipSuccess) return e;
    return pbe::enqueue_gtheta_reduce(rows, gtheta, p.n, stream);
}

extern "C" size_t pbe_adjoint_workspace_bytes(int n, int steps, int checkpoint_every)
{
    if (n <= 0 || steps <= 0 || checkpoint_every < 0) return 0;
    return pbe::adjoint_layout(n, steps, checkpoint_every).total * sizeof(double);
}

static hipError_t adjoint_checks(const pbe::RhsProblem& p, int steps, int every,
                                 const void* ws, size_t bytes)
{
    if (steps <= 0 || every < 0) return hipErrorInvalidValue;
    if (!ws) return hipErrorInvalidDevicePointer;
    if (bytes < pbe_adjoint_workspace_bytes(p.n, steps, every)) return hipErrorInvalidValue;
    return hipSuccess;
}

extern "C" hipError_t pbe_adjoint_rk4_forward(double* N, const double* x, const double* t_q,
                                              const double* bw_q,
                                              const pbe_aggregation_params* agg,
                                              const pbe_breakage_params* brk, double dt,
                                              int steps, int checkpoint_every, void* workspace,
                                              size_t workspace_bytes, hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = make_problem(p, x, t_q, bw_q, agg, brk);
    if (e != hipSuccess) return e;
    if (!N) return hipErrorInvalidDevicePointer;
    if ((e = adjoint_checks(p, steps, checkpoint_every, workspace, workspace_bytes)) != hipSuccess)
        return e;
    return pbe::enqueue_adjoint_forward(p, N, dt, steps, checkpoint_every,
                                        static_cast<double*>(workspace), stream);
}

extern "C" hipError_t pbe_adjoint_rk4_backward(const double* gN_final, double* gN0,
                                               double* gtheta, const double* x,
                                               const double* t_q, const double* bw_q,
                                               const pbe_aggregation_params* agg,
                                               const pbe_breakage_params* brk, double dt,
                                               int steps, int checkpoint_every,
                                               void* workspace, size_t workspace_bytes,
                                               hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = make_problem(p, x, t_q, bw_q, agg, brk);
    if (e != hipSuccess) return e;
    if (!gN_final || !gN0 || !gtheta) return hipErrorInvalidDevicePointer;
    if ((e = adjoint_checks(p, steps, checkpoint_every, workspace, workspace_bytes)) != hipSuccess)
        return e;
    return pbe::enqueue_adjoint_backward(p, gN_final, gN0, gtheta, dt, steps, checkpoint_every,
                                         static_cast<double*>(workspace), stream);
}

// ------------------------------------------- batched VJP and RK4 adjoint ----
extern "C" hipError_t pbe_batch_rhs_vjp(pbe_batch_plan plan, const double* N, const double* x,
                                        const double* t_q, const double* bw_q,
                                        const double* lam, double* gN, double* gtheta,
                                        void* workspace, size_t workspace_bytes,
                                        hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = batch_problem(p, plan, x, t_q, bw_q);
    if (e != hipSuccess) return e;
    if (pbe::batch_plan_needs_table(plan) && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    if (!N || !lam || !gN || !gtheta || !workspace) return hipErrorInvalidDevicePointer;
    if (workspace_bytes < pbe_vjp_workspace_bytes(p.n)) return hipErrorInvalidValue;
    double* tab = static_cast<double*>(workspace);
    double* rows = tab + 4 * (size_t)p.n;
    double* part = rows + 5 * (size_t)p.n;
    if ((e = pbe::enqueue_vjp_table(x, tab, p.n, stream)) != hipSuccess) return e;
    if ((e = pbe::enqueue_fill_zero(rows, 5 * p.n, stream)) != hipSuccess) return e;
    if ((e = pbe::enqueue_vjp(p, N, lam, tab, gN, rows, part, stream)) != hipSuccess) return e;
    return pbe::enqueue_gtheta_reduce(rows, gtheta, p.n, stream);
}

extern "C" hipError_t pbe_batch_adjoint_rk4_forward(pbe_batch_plan plan, double* N,
                                                    const double* x, const double* t_q,
                                                    const double* bw_q, double dt, int steps,
                                                    int checkpoint_every, void* workspace,
                                                    size_t workspace_bytes, hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = batch_problem(p, plan, x, t_q, bw_q);
    if (e != hipSuccess) return e;
    if (pbe::batch_plan_needs_table(plan) && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    if (!N) return hipErrorInvalidDevicePointer;
    if ((e = adjoint_checks(p, steps, checkpoint_every, workspace, workspace_bytes)) != hipSuccess)
        return e;
    return pbe::enqueue_adjoint_forward(p, N, dt, steps, checkpoint_every,
                                        static_cast<double*>(workspace), stream);
}

extern "C" hipError_t pbe_batch_adjoint_rk4_backward(pbe_batch_plan plan,
                                                     const double* gN_final, double* gN0,
                                                     double* gtheta, const double* x,
                                                     const double* t_q, const double* bw_q,
                                                     double dt, int steps, int checkpoint_every,
                                                     void* workspace, size_t workspace_bytes,
                                                     hipStream_t stream)
{
    pbe::RhsProblem p;
    hipError_t e = batch_problem(p, plan, x, t_q, bw_q);
    if (e != hipSuccess) return e;
    if (pbe::batch_plan_needs_table(plan) && (!t_q || !bw_q)) return hipErrorInvalidDevicePointer;
    if (!gN_final || !gN0 || !gtheta) return hipErrorInvalidDevicePointer;
    if ((e = adjoint_checks(p, steps, checkpoint_every, workspace, workspace_bytes)) != hipSuccess)
        return e;
    return pbe::enqueue_adjoint_backward(p, gN_final, gN0, gtheta, dt, steps, checkpoint_every,
                                         static_cast<double*>(workspace), stream);
}
