// Linear-selection breakage into two equal halves from a monodisperse start,
// the counterpart of the reference's
// examples/breakage_linear_symmetricbinary/main.cpp (docs/examples.md,
// example 4): 512 geometric bins r = 1.122018 from x_0 = 1e-18 m^3,
// N_0 = 1e14 m^-3 in bin 384, S(v) = S0 v / v_ref with S0 = 1e-3 1/s and
// v_ref = x_384, daughter table t = {1/2}, bw = {2}, explicit Euler with
// dt = 0.01 s to t = 100 s.  Checks N_tot(t) = N_0 (1 + S0 t) and the total
// volume.  With linear selection N_tot grows exactly linearly under Euler,
// so both errors are at rounding level.
#include "../common/pbe_examples.hpp"

using namespace pbe_ex;

int main()
{
    constexpr int n = 512, steps = 10000, n_print = 1000;
    constexpr double N0 = 1e14, S0 = 1e-3, dt = 0.01;
    std::printf("Linear-selection symmetric binary breakage, %d bins, explicit Euler dt = %g s\n",
                n, dt);
    print_gpu();
    const auto x = geometric_grid(n, 1e-18, 1.122018);
    std::vector<double> N(n, 0.0);
    N[384] = N0;
    const double V0 = M1(N, x);

    pbe_cuda::BreakageParams bp;
    bp.n = n;
    bp.n_quad = 1;
    bp.selection = pbe_cuda::BreakageSelection::Linear;
    bp.S0 = S0;
    bp.v_ref = x[384];
    const pbe_breakage_params cbp = c_params(bp);

    DeviceBuffer<double> dx(x), dN(N), dt_q(std::vector<double>{0.5}),
        dbw_q(std::vector<double>{2.0});
    Integrator it(dN, dx, dt_q.p, dbw_q.p, nullptr, &cbp, dt, PBE_METHOD_EULER);

    std::printf("%-8s %-14s %-14s %-10s %-10s\n", "t [s]", "N_tot", "exact", "N err", "V err");
    const double t0 = now_s();
    for (int s = 0; s <= steps; s += n_print) {
        const auto h = dN.download();
        const double t = s * dt;
        const double ex = linear_breakage_M0(N0, x[384], S0, x[384], t);
        std::printf("%-8.1f %-14.6e %-14.6e %-10.3e %-10.3e\n", t, M0(h), ex, rel_err(M0(h), ex),
                    rel_err(M1(h, x), V0));
        if (s < steps) it.step(n_print);
    }
    PBE_EX_CHECK(hipDeviceSynchronize());
    const double wall = now_s() - t0;
    const auto h = dN.download();
    const double eN = rel_err(M0(h), linear_breakage_M0(N0, x[384], S0, x[384], steps * dt));
    const double eV = rel_err(M1(h, x), V0);
    std::printf("\n  N_tot error : %.3e\n  V_tot error : %.3e\n  wall time   : %.3f s\n", eN, eV,
                wall);
    std::printf("{\"example\": \"breakage_linear_symmetricbinary\", \"N_tot\": %.17g, "
                "\"V_tot\": %.17g, \"wall_s\": %.6f}\n", eN, eV, wall);
    return 0;
}
