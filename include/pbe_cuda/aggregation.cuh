// aggregation.cuh -- C++ drop-in for pbe_cuda::launch_aggregation_rhs on HIP.
//
// Same namespace, enum, struct (fields, order, defaults) and function name as
// include/pbe_cuda/aggregation.cuh:32-120 of acuoci/pbe-fixed-pivot-cuda, with
// hipStream_t / hipError_t in place of the CUDA types.  The function is an
// inline forwarder to the C ABI (pbe_hip/pbe_hip.h).
#pragma once

#include <hip/hip_runtime_api.h>

#include "pbe_hip/pbe_hip.h"

namespace pbe_cuda {

enum class AggregationKernel : int {
    Constant      = 0,
    Sum           = 1,
    Product       = 2,
    Brownian      = 3,
    Shear         = 4,
    BrownianShear = 5
};

struct AggregationParams {
    int    n          = 0;
    double log_x0     = 0.0;
    double inv_log_r  = 0.0;
    AggregationKernel kernel_type = AggregationKernel::Constant;
    double beta0      = 1.0;
    double beta_br    = 0.0;
    double beta_sh    = 0.0;
    int    block_size = 256;
};

static_assert(sizeof(AggregationParams) == sizeof(pbe_aggregation_params),
              "AggregationParams must mirror pbe_aggregation_params");

// rhs += aggregation source; rhs must be zeroed by the caller beforehand.
inline hipError_t launch_aggregation_rhs(const double* N, const double* x, double* rhs,
                                         const AggregationParams& params,
                                         hipStream_t stream = 0)
{
    return pbe_launch_aggregation_rhs(
        N, x, rhs, reinterpret_cast<const pbe_aggregation_params*>(&params), stream);
}

}  // namespace pbe_cuda
