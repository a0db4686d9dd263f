// Linear-selection breakage with a uniform daughter distribution from an
// exponential start (Ziff and McGrady 1985), the counterpart of the
// reference's examples/breakage_linear_uniform/main.cpp (docs/examples.md,
// example 5): 200 geometric bins spanning [1e-4, 10], n(v,0) = exp(-v),
// S(v) = S0 v / v_ref with S0 = v_ref = 1, b(v|v') = 2/v' as the 8-point
// Gauss-Legendre table t = nodes on (0, 1), bw = 2 x weights, explicit Euler
// with dt = 1e-3 to t = 5.  Checks M0(t) = N_0 (1 + S0 t) = 6 and M1 = 1.
// Documented: M0 error ~2.2e-4.
#include "../common/pbe_examples.hpp"

using namespace pbe_ex;

int main()
{
    constexpr int n = 200, steps = 5000, n_print = 500, nq = 8;
    constexpr double dt = 1e-3;
    std::printf("Ziff-McGrady breakage (linear selection, uniform daughter), %d bins, "
                "explicit Euler dt = %g\n", n, dt);
    print_gpu();
    const auto x = geometric_grid_range(n, 1e-4, 1e1);
    const auto N = exponential_ic(x, 1.0, 1.0);
    const double M1_0 = M1(N, x);
    std::vector<double> t_q, w_q;
    gauss_legendre(nq, 0.0, 1.0, t_q, w_q);
    std::vector<double> bw_q(nq);
    for (int q = 0; q < nq; ++q) bw_q[q] = 2.0 * w_q[q];

    pbe_cuda::BreakageParams bp;
    bp.n = n;
    bp.n_quad = nq;
    bp.selection = pbe_cuda::BreakageSelection::Linear;
    bp.S0 = 1.0;
    bp.v_ref = 1.0;
    const pbe_breakage_params cbp = c_params(bp);

    DeviceBuffer<double> dx(x), dN(N), dt_q(t_q), dbw_q(bw_q);
    Integrator it(dN, dx, dt_q.p, dbw_q.p, nullptr, &cbp, dt, PBE_METHOD_EULER);

    std::printf("%-8s %-14s %-14s %-10s %-10s\n", "t", "M0", "exact", "M0 err", "M1 err");
    const double t0 = now_s();
    for (int s = 0; s <= steps; s += n_print) {
        const auto h = dN.download();
        const double t = s * dt;
        const double ex = linear_breakage_M0(1.0, 1.0, 1.0, 1.0, t);
        std::printf("%-8.2f %-14.6e %-14.6e %-10.3e %-10.3e\n", t, M0(h), ex, rel_err(M0(h), ex),
                    rel_err(M1(h, x), 1.0));
        if (s < steps) it.step(n_print);
    }
    PBE_EX_CHECK(hipDeviceSynchronize());
    const double wall = now_s() - t0;
    const auto h = dN.download();
    const double e0 = rel_err(M0(h), linear_breakage_M0(1.0, 1.0, 1.0, 1.0, steps * dt));
    const double e1 = rel_err(M1(h, x), 1.0);
    const double drift = rel_err(M1(h, x), M1_0);
    std::printf("\n  M0 error : %.3e\n  M1 error : %.3e (drift from t = 0: %.3e)\n"
                "  wall time: %.3f s\n", e0, e1, drift, wall);
    std::printf("{\"example\": \"breakage_linear_uniform\", \"M0\": %.17g, \"M1\": %.17g, "
                "\"M1_drift\": %.17g, \"wall_s\": %.6f}\n", e0, e1, drift, wall);
    return 0;
}
