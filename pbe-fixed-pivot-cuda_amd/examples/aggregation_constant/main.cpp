// Constant-kernel aggregation from a monodisperse start (Smoluchowski), the
// counterpart of the reference's examples/aggregation_constant/main.cpp
// (docs/examples.md, example 1): 256 geometric bins r = 1.122018 from
// x_0 = 1e-18 m^3, N_0 = 1e14 m^-3 in bin 0, beta0 = 1e-17 m^3/s, explicit
// Euler with dt = 1 s to t = 2000 s.  Checks N_tot(t) = N_0 / (1 + t/t_half),
// t_half = 2 / (beta0 N_0), and the conservation of M1.
// Documented: N_tot error ~1.7e-4, M1 error at rounding level.
#include "../common/pbe_examples.hpp"

using namespace pbe_ex;

int main()
{
    constexpr int n = 256, steps = 2000, n_print = 200;
    constexpr double N0 = 1e14, b0 = 1e-17, dt = 1.0;
    std::printf("Constant-kernel aggregation (Smoluchowski), %d bins, explicit Euler dt = %g s\n",
                n, dt);
    print_gpu();
    const auto x = geometric_grid(n, 1e-18, 1.122018);
    std::vector<double> N(n, 0.0);
    N[0] = N0;
    const double M1_0 = M1(N, x);

    pbe_cuda::AggregationParams ap;
    ap.n = n;
    grid_lookup(x, ap.log_x0, ap.inv_log_r);
    ap.kernel_type = pbe_cuda::AggregationKernel::Constant;
    ap.beta0 = b0;
    const pbe_aggregation_params cap = c_params(ap);

    DeviceBuffer<double> dx(x), dN(N);
    Integrator it(dN, dx, nullptr, nullptr, &cap, nullptr, dt, PBE_METHOD_EULER);

    std::printf("%-8s %-14s %-14s %-10s\n", "t [s]", "N_tot", "exact", "rel err");
    const double t0 = now_s();
    for (int s = 0; s <= steps; s += n_print) {
        const auto h = dN.download();
        const double t = s * dt;
        std::printf("%-8.0f %-14.6e %-14.6e %-10.3e\n", t, M0(h), smoluchowski_Ntot(N0, b0, t),
                    rel_err(M0(h), smoluchowski_Ntot(N0, b0, t)));
        if (s < steps) it.step(n_print);
    }
    PBE_EX_CHECK(hipDeviceSynchronize());
    const double wall = now_s() - t0;
    const auto h = dN.download();
    const double e_n = rel_err(M0(h), smoluchowski_Ntot(N0, b0, steps * dt));
    const double e_m1 = rel_err(M1(h, x), M1_0);
    std::printf("\n  N_tot error : %.3e\n  M1 error    : %.3e\n  wall time   : %.3f s\n", e_n,
                e_m1, wall);
    std::printf("{\"example\": \"aggregation_constant\", \"N_tot\": %.17g, \"M1\": %.17g, "
                "\"wall_s\": %.6f}\n", e_n, e_m1, wall);
    return 0;
}
