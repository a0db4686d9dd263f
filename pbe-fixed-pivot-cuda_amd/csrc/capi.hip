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
// the autotuned table (tile_table.inc: rows {n_max, kernel_type, tile_u,
// band_split}, generated on an MI355X by pbe-fixed-pivot-cuda_amd/autotune.py;
// kernel_type -1 matches any).
struct TileRule { int n_max, kernel, u, w; };
static const TileRule kTileTable[] = {
#include "tile_table.inc"
};

static const TileRule& tile_rule(int n, int kernel_type)
{
    static const TileRule fallback{1 << 30, -1, 4, 4};
    for (const TileRule& r : kTileTable)
        if (n <= r.n_max && (r.kernel < 0 || r.kernel == kernel_type)) return r;
    return fallback;
}

int default_tile_u(int n, int kernel_type) { return tile_rule(n, kernel_type).u; }

int default_band_split(int n, int kernel_type) { return tile_rule(n, kernel_type).w; }

// The table's columns per lane, widened on fine geometric grids.  A strip
// pair (j, k) lies more than T = 64 U bins off the diagonal, and it is far
// (both births in bins k, k+1) when x_j < (r - 1) x_k, i.e. k - j > log_r
// (1 / (r - 1)) bins; a strip chunk holding a nearer pair takes the per-pair
// path for all its pairs, and those chunks sit together in one or two
// workgroups.  So from n = 6000 up, a grid whose reach exceeds T but not 512
// bins takes U = 8, which leaves the strips far pairs only; a finer grid
// takes U = 8 from n = 32768 up (fewer, wider near chunks).  Measured, one
// aggregation call, U = 8 against U = 4: n = 6000, r = 1.01 (reach 463 bins)
// 0.067 / 0.107 ms; n = 8192, r = 1.01 0.070 / 0.184 ms; n = 16384, r = 1.01
// 0.12 / 0.40 ms; n = 49152, r = 1.006 (reach 866) 2.58 / 3.81 ms.  Below
// n = 6000 the fused launch at U = 4 is faster (n = 4096, r = 1.01: 0.061
// against 0.138 ms), as is U = 4 for the finer grid at n = 16384 (r = 1.005:
// 0.53 against 0.57 ms).  General grids keep the table's choice.
int grid_tile_u(const pbe_aggregation_params& p)
{
    const int u = default_tile_u(p.n, p.kernel_type);
    const double il = p.inv_log_r;
    if (p.n < 6000 || !(il > 0.0) || !std::isfinite(il) || u == 8) return u;
    const double reach = -std::log(std::expm1(1.0 / il)) * il;   // log_r (1 / (r - 1))
    if (reach <= 64.0 * u) return u;
    return (reach <= 512.0 || p.n >= 32768) ? 8 : u;
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
    hipError_t e = pbe::launch_aggregation(N, x, rhs, *p,
                                           pbe::grid_tile_u(*p),
                                           pbe::default_band_split(p->n, p->kernel_type),
                                           stream);
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
    int u = opts ? opts->tile_u : 0;
    int w = opts ? opts->band_split : 0;
    if (u == 0) u = pbe::grid_tile_u(*p);
    if (w == 0) w = pbe::default_band_split(p->n, p->kernel_type);
    if (u != 1 && u != 2 && u != 4 && u != 8) return hipErrorInvalidValue;
    if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return hipErrorInvalidValue;
    hipError_t e = pbe::launch_aggregation(N, x, rhs, *p, u, w, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

extern "C" int pbe_aggregation_default_tile(int n, int kernel_type)
{
    return pbe::default_tile_u(n, kernel_type);
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
    return pbe::batch_plan_create(out, instances, count, kernel_type, selection, n_quad);
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
