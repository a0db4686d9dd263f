// Two-stage flocculation with aggregation and breakage, the counterpart of
// the reference's examples/flocculation_two_stage/main.cpp (the case study
// of its paper, Section 6; constants and dimensionless form as at
// main.cpp:97-205): silica particles d0 = 0.5 um at volume fraction 5e-5 in
// water at 298 K; Brownian + shear aggregation; power-law selection
// S = k_br G / G_ref (v / v_crit)^2 with an erosion daughter (fragments
// 0.95 v and 0.05 v); shear rate G = 100 1/s up to t_mid = 600 s, then
// 20 1/s to t_end = 1800 s; RK4 with dt = 0.5 s; N = 960 geometric sections
// over d in [0.05, 200] um; a log-normal start centred on d0 (sigma_ln 0.3).
// Variables are dimensionless: x = v / v0, N / N0, t / t_c with
// t_c = 1 / (beta0 N0).
//
// The time loop runs on the device (one RK4 step per hipGraph replay, a graph
// per shear rate) and stops every 60 steps to print d43, M0 and the drift of
// M1, as the reference does.  As there, the shear rate of a step is the
// high one while the accumulated dimensionless time times t_c is below t_mid.
//
//   example_flocculation_two_stage [N [t_end_s [t_mid_s [final_state_file]]]]
#include "../common/pbe_examples.hpp"

#include <fstream>

using namespace pbe_ex;

namespace {

constexpr double kB = 1.380649e-23, kPi = 3.14159265358979323846;
constexpr double d0 = 0.5e-6, phi = 5.0e-5, T = 298.0, mu = 1.0e-3;
constexpr double d_min = 0.05e-6, d_max = 200.0e-6;
constexpr double G_high = 100.0, G_low = 20.0, G_ref = 1.0, dt_s = 0.5;
constexpr double d_crit = 5.0e-6, k_br = 1.0e-3, alpha_br = 2.0, eps_ero = 0.05;
constexpr int n_print = 60;

struct Scales {
    double v0, N0, beta_br, beta0, t_c;
};

Scales scales()
{
    Scales s;
    s.v0 = (kPi / 6.0) * d0 * d0 * d0;
    s.N0 = phi / s.v0;
    s.beta_br = 2.0 * kB * T / (3.0 * mu);
    s.beta0 = 4.0 * s.beta_br;
    s.t_c = 1.0 / (s.beta0 * s.N0);
    return s;
}

pbe_cuda::AggregationParams agg_params(int n, const std::vector<double>& x, double G,
                                       const Scales& s)
{
    pbe_cuda::AggregationParams p;
    p.n = n;
    grid_lookup(x, p.log_x0, p.inv_log_r);
    p.kernel_type = pbe_cuda::AggregationKernel::BrownianShear;
    p.beta0 = 0.0;
    p.beta_br = 0.25;                           // 2 kB T / (3 mu) / beta0
    p.beta_sh = G * s.v0 / (kPi * s.beta0);     // (4/3) G (3 v0 / 4 pi) / beta0
    return p;
}

pbe_cuda::BreakageParams brk_params(int n, double G, const Scales& s)
{
    pbe_cuda::BreakageParams p;
    p.n = n;
    p.n_quad = 2;
    p.selection = pbe_cuda::BreakageSelection::PowerLaw;
    p.S0 = k_br * G / G_ref * s.t_c;
    p.v_ref = std::pow(d_crit / d0, 3.0);
    p.alpha = alpha_br;
    p.v_min = 0.0;
    return p;
}

}  // namespace

int main(int argc, char** argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 960;
    const double t_end_s = argc > 2 ? std::atof(argv[2]) : 1800.0;
    const double t_mid_s = argc > 3 ? std::atof(argv[3]) : 600.0;
    const char* state_file = argc > 4 ? argv[4] : nullptr;
    const Scales sc = scales();
    std::printf("Two-stage flocculation (dimensionless form), N = %d, RK4 dt = %.1f s, "
                "t_end = %.0f s, G %g -> %g 1/s at %.0f s\n", n, dt_s, t_end_s, G_high, G_low,
                t_mid_s);
    print_gpu();
    std::printf("t_c = %.4f s, beta_sh_nd = %.4f (high) / %.4f (low), S0_nd = %.4f / %.4f\n\n",
                sc.t_c, G_high * sc.v0 / (kPi * sc.beta0), G_low * sc.v0 / (kPi * sc.beta0),
                k_br * G_high * sc.t_c, k_br * G_low * sc.t_c);

    const auto x = geometric_grid_range(n, std::pow(d_min / d0, 3.0), std::pow(d_max / d0, 3.0));
    const auto N = lognormal_ic_nd(x, 0.3);
    const double M1_ref = M1(N, x);

    // the step count and the switch step, by the same time accumulation as
    // the reference's loop (the switch falls one step after t_mid here)
    const double dt_nd = dt_s / sc.t_c;
    const int steps = static_cast<int>(t_end_s / dt_s);
    int n_high = 0;
    {
        double t_nd = 0.0;
        for (int s = 0; s < steps; ++s) {
            if (t_nd * sc.t_c < t_mid_s) n_high = s + 1;
            t_nd += dt_nd;
        }
    }

    const auto a_hi = c_params(agg_params(n, x, G_high, sc));
    const auto a_lo = c_params(agg_params(n, x, G_low, sc));
    const auto b_hi = c_params(brk_params(n, G_high, sc));
    const auto b_lo = c_params(brk_params(n, G_low, sc));
    DeviceBuffer<double> dx(x), dN(N), dt_q(std::vector<double>{1.0 - eps_ero, eps_ero}),
        dbw_q(std::vector<double>{1.0, 1.0});
    Integrator high(dN, dx, dt_q.p, dbw_q.p, &a_hi, &b_hi, dt_nd, PBE_METHOD_RK4);
    Integrator low(dN, dx, dt_q.p, dbw_q.p, &a_lo, &b_lo, dt_nd, PBE_METHOD_RK4);

    std::printf("%-8s %-10s %-12s %-12s %-10s %-8s\n", "t [s]", "d43 [um]", "M0_nd", "M1_nd",
                "err M1", "stage");
    const int mid_step = static_cast<int>(t_mid_s / dt_s + 0.5);   // the state at t = t_mid
    double d43_mid = 0.0;
    const double t0 = now_s();
    for (int s = 0; s <= steps;) {
        const auto h = dN.download();
        const double d = d43(h, x, d0) * 1e6;
        if (s == mid_step) d43_mid = d;
        if (s % n_print == 0 || s == steps)
            std::printf("%-8.1f %-10.3f %-12.4e %-12.4e %-10.3e %-8s\n", s * dt_s, d, M0(h),
                        M1(h, x), rel_err(M1(h, x), M1_ref), s < n_high ? "High-G" : "Low-G");
        if (s == steps) break;
        // run to the next print or to t_mid, whichever is first; each step
        // takes the shear rate of its own start time
        int next = std::min(steps, (s / n_print + 1) * n_print);
        if (s < mid_step && next > mid_step) next = mid_step;
        for (int k = s; k < next; ++k) (k < n_high ? high : low).step(1);
        s = next;
    }
    PBE_EX_CHECK(hipDeviceSynchronize());
    const double wall = now_s() - t0;
    const auto h = dN.download();
    const double d43_end = d43(h, x, d0) * 1e6, err_M1 = rel_err(M1(h, x), M1_ref);
    std::printf("\nFinal t = %.0f s (%d high-G steps):\n  d43 at t_mid : %.4f um\n"
                "  d43 at t_end : %.4f um\n  M0_nd        : %.6e\n  M1 drift     : %.3e\n"
                "  wall time    : %.3f s (%d RK4 steps, %d sections)\n", t_end_s, n_high, d43_mid,
                d43_end, M0(h), err_M1, wall, steps, n);
    if (state_file) {
        std::ofstream f(state_file, std::ios::binary);
        f.write(reinterpret_cast<const char*>(h.data()), h.size() * sizeof(double));
    }
    std::printf("{\"example\": \"flocculation_two_stage\", \"n\": %d, \"steps\": %d, "
                "\"steps_high\": %d, \"d43_mid_um\": %.17g, \"d43_end_um\": %.17g, "
                "\"M0_end\": %.17g, \"M1_drift\": %.17g, \"wall_s\": %.6f}\n", n, steps, n_high,
                d43_mid, d43_end, M0(h), err_M1, wall);
    return 0;
}
