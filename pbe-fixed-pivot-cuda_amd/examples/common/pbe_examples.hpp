// pbe_examples.hpp -- helpers shared by the example programs: grids, initial
// conditions, moments and d43, the analytic moment laws the examples check
// against, device buffers, and a device time integrator (hipGraph replays of
// pbe_integrator_graph_create, so a step makes no host round trip).
//
// The grids and initial conditions follow the conventions of the
// reference's example harness (examples/common/pbe_examples_utils.hpp:
// make_geometric_grid :83-90 by repeated multiplication,
// make_geometric_grid_range :93-99, make_exponential_ic :121-136 as exact
// cell integrals, make_lognormal_ic_nd :142-158); the analytic laws are those
// of :198-296.  Everything here is host code over the C ABI and the drop-in
// C++ headers.
#pragma once

#include <pbe_cuda/pbe_cuda.cuh>
#include <pbe_hip/pbe_hip.h>
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace pbe_ex {

#define PBE_EX_CHECK(call)                                                        \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s:%d: %s failed: %s\n", __FILE__, __LINE__, #call, \
                         hipGetErrorString(e_));                                  \
            std::exit(EXIT_FAILURE);                                              \
        }                                                                         \
    } while (0)

// ------------------------------------------------------------------ grids --
inline std::vector<double> geometric_grid(int n, double x0, double r)
{
    std::vector<double> x(n);
    x[0] = x0;
    for (int i = 1; i < n; ++i) x[i] = x[i - 1] * r;
    return x;
}

inline std::vector<double> geometric_grid_range(int n, double v_lo, double v_hi)
{
    return geometric_grid(n, v_lo, std::pow(v_hi / v_lo, 1.0 / (n - 1)));
}

// log_x0 and inv_log_r of a geometric grid, as the library's callers pass them
inline void grid_lookup(const std::vector<double>& x, double& log_x0, double& inv_log_r)
{
    log_x0 = std::log(x[0]);
    inv_log_r = 1.0 / std::log(x[1] / x[0]);
}

// ------------------------------------------------------ initial conditions --
// n(v) = (N0/vc) exp(-v/vc) integrated exactly over the cells
// [x_i / sqrt(r), x_i sqrt(r)]
inline std::vector<double> exponential_ic(const std::vector<double>& x, double N0, double vc)
{
    const double sr = std::sqrt(x[1] / x[0]);
    std::vector<double> N(x.size());
    for (size_t i = 0; i < x.size(); ++i)
        N[i] = N0 * (std::exp(-(x[i] / sr) / vc) - std::exp(-(x[i] * sr) / vc));
    return N;
}

// exp(-ln(x)^2 / (2 sigma^2)) x ln r, normalised to a sum of 1 (centred at x = 1)
inline std::vector<double> lognormal_ic_nd(const std::vector<double>& x, double sigma)
{
    const double lr = std::log(x[1] / x[0]);
    std::vector<double> N(x.size());
    double s = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
        const double z = std::log(x[i]) / sigma;
        N[i] = std::exp(-0.5 * z * z) * (x[i] * lr);
        s += N[i];
    }
    for (double& v : N) v /= s;
    return N;
}

// m-point Gauss-Legendre nodes and weights on (a, b): Newton iteration on the
// Legendre polynomial P_m from the Chebyshev-like guess cos(pi (i - 1/4) /
// (m + 1/2)), weights 2 / ((1 - t^2) P_m'(t)^2).
inline void gauss_legendre(int m, double a, double b, std::vector<double>& t,
                           std::vector<double>& w)
{
    t.assign(m, 0.0);
    w.assign(m, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (m + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (m + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;   // P_0, P_1
            for (int k = 2; k <= m; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = m * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        t[i] = 0.5 * (a + b) - 0.5 * (b - a) * z;
        t[m - 1 - i] = 0.5 * (a + b) + 0.5 * (b - a) * z;
        w[i] = w[m - 1 - i] = 0.5 * (b - a) * wi;
    }
}

// ---------------------------------------------------------------- moments --
inline double moment(const std::vector<double>& N, const std::vector<double>& x, int k)
{
    double m = 0.0;
    for (size_t i = 0; i < N.size(); ++i) {
        double xk = 1.0;
        for (int j = 0; j < k; ++j) xk *= x[i];
        m += N[i] * xk;
    }
    return m;
}

inline double M0(const std::vector<double>& N)
{
    double s = 0.0;
    for (double v : N) s += v;
    return s;
}

inline double M1(const std::vector<double>& N, const std::vector<double>& x) { return moment(N, x, 1); }
inline double M2(const std::vector<double>& N, const std::vector<double>& x) { return moment(N, x, 2); }

// volume-weighted mean diameter d43 = sum N d^4 / sum N d^3 for pivots that
// are volumes x in units of a reference volume of diameter d_ref
inline double d43(const std::vector<double>& N, const std::vector<double>& x, double d_ref)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < N.size(); ++i) {
        num += N[i] * std::pow(x[i], 4.0 / 3.0);
        den += N[i] * x[i];
    }
    return den > 0.0 ? d_ref * num / den : 0.0;
}

inline double rel_err(double a, double b) { return std::fabs(a - b) / std::fabs(b); }

// ------------------------------------------------------- analytic moments --
// constant kernel, monodisperse start: N_tot = N0 / (1 + t / t_half), t_half = 2 / (b0 N0)
inline double smoluchowski_Ntot(double N0, double b0, double t)
{
    return N0 / (1.0 + t / (2.0 / (b0 * N0)));
}
// sum kernel, exponential start (Golovin): M0 = N0 exp(-b0 N0 vc t)
inline double golovin_M0(double N0, double vc, double b0, double t)
{
    return N0 * std::exp(-b0 * N0 * vc * t);
}
// product kernel, exponential start: M0 = N0 (1 - tau/2), M2 = 2 N0 vc^2 / (1 - b0 M2(0) t)
inline double product_M0(double N0, double vc, double b0, double t)
{
    return N0 * (1.0 - 0.5 * b0 * N0 * vc * vc * t);
}
inline double product_M2(double N0, double vc, double b0, double t)
{
    const double m20 = 2.0 * N0 * vc * vc;
    return m20 / (1.0 - b0 * m20 * t);
}
// linear selection (any daughter, volume-conserving): M0 = N0 (1 + S0 vc t / v_ref)
inline double linear_breakage_M0(double N0, double vc, double S0, double v_ref, double t)
{
    return N0 * (1.0 + S0 * vc / v_ref * t);
}

// ---------------------------------------------------------- device memory --
template <class T>
struct DeviceBuffer {
    T* p = nullptr;
    size_t n = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t count) : n(count) { PBE_EX_CHECK(hipMalloc(&p, n * sizeof(T))); }
    explicit DeviceBuffer(const std::vector<T>& h) : DeviceBuffer(h.size()) { upload(h); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { if (p) (void)hipFree(p); }
    void upload(const std::vector<T>& h)
    {
        PBE_EX_CHECK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
    }
    std::vector<T> download() const
    {
        std::vector<T> h(n);
        PBE_EX_CHECK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
};

// ------------------------------------------------------ device integrator --
// RK4 or Euler steps of dN/dt = R_agg + R_brk on the device: one step is
// captured once in a hipGraph and replayed (agg or brk may be null).
struct Integrator {
    pbe_graph graph = nullptr;
    DeviceBuffer<double> work;
    Integrator(DeviceBuffer<double>& N, const DeviceBuffer<double>& x, const double* t_q,
               const double* bw_q, const pbe_aggregation_params* agg,
               const pbe_breakage_params* brk, double dt, int method)
        : work(pbe_integrator_workspace_bytes((int)N.n, method) / sizeof(double))
    {
        PBE_EX_CHECK(pbe_integrator_graph_create(&graph, N.p, x.p, t_q, bw_q, agg, brk, dt, 1,
                                                 method, work.p, work.n * sizeof(double)));
    }
    ~Integrator() { if (graph) (void)pbe_graph_destroy(graph); }
    void step(int count, hipStream_t s = nullptr)
    {
        for (int i = 0; i < count; ++i) PBE_EX_CHECK(pbe_graph_launch(graph, s));
    }
};

// C-ABI parameter blocks from the drop-in C++ structs
inline pbe_aggregation_params c_params(const pbe_cuda::AggregationParams& p)
{
    pbe_aggregation_params c;
    c.n = p.n;
    c.log_x0 = p.log_x0;
    c.inv_log_r = p.inv_log_r;
    c.kernel_type = static_cast<int>(p.kernel_type);
    c.beta0 = p.beta0;
    c.beta_br = p.beta_br;
    c.beta_sh = p.beta_sh;
    c.block_size = p.block_size;
    return c;
}

inline pbe_breakage_params c_params(const pbe_cuda::BreakageParams& p)
{
    pbe_breakage_params c;
    c.n = p.n;
    c.n_quad = p.n_quad;
    c.selection = static_cast<int>(p.selection);
    c.S0 = p.S0;
    c.v_ref = p.v_ref;
    c.alpha = p.alpha;
    c.v_min = p.v_min;
    c.block_size = p.block_size;
    return c;
}

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

inline void print_gpu()
{
    hipDeviceProp_t prop{};
    PBE_EX_CHECK(hipGetDeviceProperties(&prop, 0));
    std::printf("GPU : %s (%s, %.0f MB)\n\n", prop.name, prop.gcnArchName,
                prop.totalGlobalMem / 1.0e6);
}

}  // namespace pbe_ex
