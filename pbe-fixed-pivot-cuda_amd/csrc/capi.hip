// capi.hip -- extern "C" boundary: argument checks, dispatch, launch status.
//
// Check order and error codes follow src/aggregation.cu:23-24,54 and
// src/breakage.cu:23-24,46,74 of the reference (n, then pointers, then the
// enum).  Every check runs before the first HIP runtime call.
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

int default_tile_u(int n, int /*kernel_type*/)
{
    return n <= 512 ? 2 : 4;
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
                                           pbe::default_tile_u(p->n, p->kernel_type),
                                           stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
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

extern "C" const char* pbe_version(void)
{
    return "pbe_hip 0.1.0 gfx950";
}
