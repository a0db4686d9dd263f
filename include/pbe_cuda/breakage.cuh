// breakage.cuh -- C++ drop-in for pbe_cuda::launch_breakage_rhs on HIP.
//
// Same namespace, enum, struct (fields, order, defaults) and function name as
// include/pbe_cuda/breakage.cuh:30-129 of acuoci/pbe-fixed-pivot-cuda, with
// hipStream_t / hipError_t in place of the CUDA types.  The function is an
// inline forwarder to the C ABI (pbe_hip/pbe_hip.h).
#pragma once

#include <hip/hip_runtime_api.h>

#include "pbe_hip/pbe_hip.h"

namespace pbe_cuda {

enum class BreakageSelection : int {
    Constant  = 0,
    Linear    = 1,
    PowerLaw  = 2,
    Threshold = 3
};

struct BreakageParams {
    int    n      = 0;
    int    n_quad = 0;
    BreakageSelection selection = BreakageSelection::Constant;
    double S0     = 1.0;
    double v_ref  = 1.0;
    double alpha  = 1.0;
    double v_min  = 0.0;
    int    block_size = 256;
};

static_assert(sizeof(BreakageParams) == sizeof(pbe_breakage_params),
              "BreakageParams must mirror pbe_breakage_params");

// rhs += breakage source (death and quadrature birth); rhs zeroed by the caller.
inline hipError_t launch_breakage_rhs(const double* N, const double* x,
                                      const double* t_q, const double* bw_q,
                                      double* rhs, const BreakageParams& params,
                                      hipStream_t stream = 0)
{
    return pbe_launch_breakage_rhs(
        N, x, t_q, bw_q, rhs, reinterpret_cast<const pbe_breakage_params*>(&params),
        stream);
}

}  // namespace pbe_cuda
