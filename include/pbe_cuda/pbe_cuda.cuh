// pbe_cuda.cuh -- umbrella header, as include/pbe_cuda/pbe_cuda.cuh:17-18 of
// the reference.
#pragma once

#include "pbe_cuda/aggregation.cuh"
#include "pbe_cuda/breakage.cuh"
