// pbe_device.hpp -- device helpers shared by the gfx950 kernels.
//
// FP64 throughout.  Atomics are the hardware FP64 adds of CDNA
// (ds_add_f64 in LDS, global_atomic_add_f64 in HBM; the library is built with
// -munsafe-fp-atomics so neither becomes a compare-and-swap loop).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbe {

constexpr int kWave = 64;

__device__ __forceinline__ void lds_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void global_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Loads typed with their address space.  A choice between an LDS element and
// a global one written as `cond ? lds[i] : glob[j]` becomes a single generic
// (flat) load, which counts against both the LDS and the vector-memory
// counters and so waits for every global atomic still in flight; with the
// address space fixed, the two stay a ds_read and a global_load.
__device__ __forceinline__ double lds_ld(const double* p)
{
    return *(const __attribute__((address_space(3))) double*)p;
}

__device__ __forceinline__ double glb_ld(const double* p)
{
    return *(const __attribute__((address_space(1))) double*)p;
}

// Issue priority of this wave among the waves of its SIMD (0 lowest, 3
// highest); p must be the same in every lane.
__device__ __forceinline__ void set_wave_priority(int p)
{
    switch (p) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(0); break;
    }
}

// Move a double across lanes with a DPP control (both 32-bit halves); lanes
// whose row is not in ROW_MASK read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave in a fixed order; every lane gets the
// total.  Data-parallel-primitive moves instead of LDS permutes: pairs
// (quad_perm 1,0,3,2), quads (2,3,0,1), half rows (row_half_mirror), rows of
// 16 (row_mirror), then row 0 into row 1 and row 2 into row 3
// (row_bcast15), rows 0-1 into row 3 (row_bcast31); lane 63 holds the total.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0xB1, 0xF>(v);
    v += dpp_move<0x4E, 0xF>(v);
    v += dpp_move<0x141, 0xF>(v);
    v += dpp_move<0x140, 0xF>(v);
    v += dpp_move<0x142, 0xA>(v);
    v += dpp_move<0x143, 0xC>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Unique right bracket hi with x[hi-1] <= v < x[hi], for x[0] < v < x[n-1]
// and strictly increasing x.  This is the bracket the reference's binary
// search returns (src/aggregation_kernels.cuh:125-133); `lo` is any index
// already known to satisfy x[lo] <= v.
__device__ __forceinline__ int bracket_search(const double* __restrict__ x,
                                              int lo, int hi, double v)
{
    while (hi - lo > 1) {
        int mid = lo + ((hi - lo) >> 1);
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    return hi;
}

// Largest i in [lo, hi) with xw[i] <= v, for xw[lo] <= v < xw[hi] (an LDS
// copy of consecutive pivots).
__device__ __forceinline__ int window_bracket_lo(const double* xw, int lo, int hi, double v)
{
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if (xw[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// Geometric-grid bracket: the reference's O(1) log-ratio guess
// (src/aggregation_kernels.cuh:121-124) followed by a fix-up against the real
// pivots, so the result is the same unique bracket as the search above even
// when log() rounding moves floor() across an integer.
__device__ __forceinline__ int bracket_geometric(const double* __restrict__ x,
                                                 int n, double v,
                                                 double log_x0, double inv_log_r)
{
    double pos = (log(v) - log_x0) * inv_log_r;
    int hi = (int)floor(pos) + 1;
    hi = hi < 1 ? 1 : (hi > n - 1 ? n - 1 : hi);
    while (hi < n - 1 && x[hi] <= v) ++hi;
    while (hi > 1 && x[hi - 1] > v) --hi;
    return hi;
}

// ---------------------------------------------------------------------------
// Separable form of the collision kernels.
//
// Every kernel of src/aggregation_kernels.cuh:52-89 is a short sum of
// products  beta(x_j, x_k) = sum_r a_r(j) * phi_r(k), all terms non-negative:
//   Constant       b0                                  R = 1
//   Sum            b0 x_j * 1 + b0 * x_k                R = 2
//   Product        b0 x_j * x_k                         R = 1
//   Brownian       b_br(2 + c_j/c_k + c_k/c_j)          R = 3   (c = cbrt x)
//   Shear          b_sh(c_j + c_k)^3 expanded           R = 4
//   BrownianShear  both, sharing phi = {1,1/c,c,c^2,c^3} R = 5
// With A_r(j) = N_j a_r(j) and B_r(k) = N_k phi_r(k) the pair rate is
// beta_jk N_j N_k = sum_r A_r(j) B_r(k): R multiply-adds per pair, no cbrt
// or divide in the pair loop.  Because every term is non-negative the
// relative rounding error is a few ulp of the rate itself.
// ---------------------------------------------------------------------------
struct KernelCoef {
    double b0, b_br, b_sh;
};

template <int KT> struct Kern;

template <> struct Kern<0> {
    static constexpr int R = 1;
    static constexpr bool kCbrt = false;
    __device__ static void row(double* a, double, double, double, const KernelCoef& c)
    { a[0] = c.b0; }
    __device__ static void col(double* f, double, double, double) { f[0] = 1.0; }
};

template <> struct Kern<1> {
    static constexpr int R = 2;
    static constexpr bool kCbrt = false;
    __device__ static void row(double* a, double x, double, double, const KernelCoef& c)
    { a[0] = c.b0 * x; a[1] = c.b0; }
    __device__ static void col(double* f, double x, double, double)
    { f[0] = 1.0; f[1] = x; }
};

template <> struct Kern<2> {
    static constexpr int R = 1;
    static constexpr bool kCbrt = false;
    __device__ static void row(double* a, double x, double, double, const KernelCoef& c)
    { a[0] = c.b0 * x; }
    __device__ static void col(double* f, double x, double, double) { f[0] = x; }
};

template <> struct Kern<3> {
    static constexpr int R = 3;
    static constexpr bool kCbrt = true;
    __device__ static void row(double* a, double, double cb, double icb, const KernelCoef& c)
    { a[0] = 2.0 * c.b_br; a[1] = c.b_br * cb; a[2] = c.b_br * icb; }
    __device__ static void col(double* f, double, double cb, double icb)
    { f[0] = 1.0; f[1] = icb; f[2] = cb; }
};

template <> struct Kern<4> {
    static constexpr int R = 4;
    static constexpr bool kCbrt = true;
    __device__ static void row(double* a, double, double cb, double, const KernelCoef& c)
    {
        a[0] = c.b_sh * (cb * cb * cb);
        a[1] = 3.0 * c.b_sh * (cb * cb);
        a[2] = 3.0 * c.b_sh * cb;
        a[3] = c.b_sh;
    }
    __device__ static void col(double* f, double, double cb, double)
    { f[0] = 1.0; f[1] = cb; f[2] = cb * cb; f[3] = cb * cb * cb; }
};

template <> struct Kern<5> {
    static constexpr int R = 5;
    static constexpr bool kCbrt = true;
    __device__ static void row(double* a, double, double cb, double icb, const KernelCoef& c)
    {
        a[0] = 2.0 * c.b_br + c.b_sh * (cb * cb * cb);
        a[1] = c.b_br * cb;
        a[2] = c.b_br * icb + 3.0 * c.b_sh * (cb * cb);
        a[3] = 3.0 * c.b_sh * cb;
        a[4] = c.b_sh;
    }
    __device__ static void col(double* f, double, double cb, double icb)
    { f[0] = 1.0; f[1] = icb; f[2] = cb; f[3] = cb * cb; f[4] = cb * cb * cb; }
};

// Cube root c = x^(1/3) and its reciprocal 1/c of a positive normal double
// without a divide or transcendental: a first guess of x^(-1/3) from the
// exponent bits (the integer image of a double is an affine function of its
// base-2 logarithm, so y ~ K - I(x)/3 with K = (4/3)(1023 - sigma) 2^52,
// sigma = 0.0450465 centring the piecewise-linear error), four Newton steps
// r <- r (4 - x r^3) / 3 for f(r) = r^-3 - x (quadratic convergence from a
// ~6 % guess), then c = x r^2 refined by one Newton step on c^3 - x.
// Measured against glibc cbrt over x in [1e-300, 1e300]: c within 3 ulp,
// 1/c within 6 ulp.  Non-positive, subnormal or non-finite x falls back to
// cbrt().
__host__ __device__ __forceinline__ void cbrt_pair(double x, double& c, double& ic)
{
    if (!(x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308)) {
        c = cbrt(x);
        ic = 1.0 / c;
        return;
    }
    const unsigned long long K = 6142639396532537344ull;   // (4/3)(1023 - 0.0450465) 2^52
    unsigned long long bits;
    __builtin_memcpy(&bits, &x, sizeof(bits));
    // I(x)/3 through a double: its rounding (~2^-53 relative) is far below the
    // guess's own ~6 % error, and it avoids a 64-bit integer division
    bits = K - (unsigned long long)((double)bits * (1.0 / 3.0));
    double r;
    __builtin_memcpy(&r, &bits, sizeof(r));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double r3 = r * r * r;
        r = r * fma(-x, r3, 4.0) * (1.0 / 3.0);
    }
    double cc = x * r * r;
    const double e = fma(cc * cc, cc, -x);
    cc = fma(-e * (1.0 / 3.0), r * r, cc);
    c = cc;
    ic = r;
}

// cbrt and its reciprocal for one pivot (only for kernels that need them).
template <int KT>
__device__ __forceinline__ void pivot_roots(double x, double& cb, double& icb)
{
    if constexpr (Kern<KT>::kCbrt) {
        cbrt_pair(x, cb, icb);
    } else {
        cb = 0.0;
        icb = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Breakage selection S(v): src/breakage_kernels.cuh:51-65, same expressions.
// ---------------------------------------------------------------------------
template <int SEL>
__device__ __forceinline__ double selection(double v, double S0, double v_ref,
                                            double alpha, double v_min)
{
    if constexpr (SEL == 0) return S0;
    else if constexpr (SEL == 1) return S0 * v / v_ref;
    else if constexpr (SEL == 2) return S0 * pow(v / v_ref, alpha);
    else return (v > v_min) ? S0 : 0.0;
}

}  // namespace pbe
