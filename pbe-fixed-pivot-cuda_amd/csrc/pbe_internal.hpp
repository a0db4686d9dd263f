// pbe_internal.hpp -- launchers shared between the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "pbe_hip/pbe_hip.h"

namespace pbe {

// Enqueue the aggregation tiles (U = columns per lane: 2, 4 or 8).
hipError_t launch_aggregation(const double* N, const double* x, double* rhs,
                              const pbe_aggregation_params& p, int tile_u,
                              hipStream_t stream);

// Enqueue the fused breakage death + birth kernel.
hipError_t launch_breakage(const double* N, const double* x, const double* t_q,
                           const double* bw_q, double* rhs,
                           const pbe_breakage_params& p, hipStream_t stream);

// Columns per lane picked for an aggregation problem of size n.
int default_tile_u(int n, int kernel_type);

}  // namespace pbe
