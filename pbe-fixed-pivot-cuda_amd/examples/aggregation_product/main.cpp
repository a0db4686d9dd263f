// Product-kernel aggregation from an exponential start, the counterpart of
// the reference's examples/aggregation_product/main.cpp (docs/examples.md,
// example 3): the grid, start and integrator of the sum-kernel example,
// beta0 = 1, to t = 0.3 (gelation at t = 0.5).  Checks M0(t) = N_0 (1 - t/2),
// M2(t) = 2 / (1 - 2t) and M1 = 1.
// Documented: M0 error ~1.7e-4, M2 error ~4.8e-3, M1 error ~1.6e-4.
#include "../common/pbe_examples.hpp"

using namespace pbe_ex;

int main()
{
    constexpr int n = 300, steps = 3000, n_print = 500;
    constexpr double dt = 1e-4;
    std::printf("Product-kernel aggregation, %d bins, RK4 dt = %g\n", n, dt);
    print_gpu();
    const auto x = geometric_grid_range(n, 1e-4, 1e4);
    const auto N = exponential_ic(x, 1.0, 1.0);

    pbe_cuda::AggregationParams ap;
    ap.n = n;
    grid_lookup(x, ap.log_x0, ap.inv_log_r);
    ap.kernel_type = pbe_cuda::AggregationKernel::Product;
    ap.beta0 = 1.0;
    const pbe_aggregation_params cap = c_params(ap);

    DeviceBuffer<double> dx(x), dN(N);
    Integrator it(dN, dx, nullptr, nullptr, &cap, nullptr, dt, PBE_METHOD_RK4);

    std::printf("%-8s %-12s %-12s %-10s %-10s %-10s\n", "t", "M0", "M2", "M0 err", "M2 err",
                "M1 err");
    const double t0 = now_s();
    for (int s = 0; s <= steps; s += n_print) {
        const auto h = dN.download();
        const double t = s * dt;
        std::printf("%-8.3f %-12.6e %-12.6e %-10.3e %-10.3e %-10.3e\n", t, M0(h), M2(h, x),
                    rel_err(M0(h), product_M0(1.0, 1.0, 1.0, t)),
                    rel_err(M2(h, x), product_M2(1.0, 1.0, 1.0, t)), rel_err(M1(h, x), 1.0));
        if (s < steps) it.step(n_print);
    }
    PBE_EX_CHECK(hipDeviceSynchronize());
    const double wall = now_s() - t0;
    const auto h = dN.download();
    const double t = steps * dt;
    const double e0 = rel_err(M0(h), product_M0(1.0, 1.0, 1.0, t));
    const double e2 = rel_err(M2(h, x), product_M2(1.0, 1.0, 1.0, t));
    const double e1 = rel_err(M1(h, x), 1.0);
    std::printf("\n  M0 error : %.3e\n  M2 error : %.3e\n  M1 error : %.3e\n  wall time: %.3f s\n",
                e0, e2, e1, wall);
    std::printf("{\"example\": \"aggregation_product\", \"M0\": %.17g, \"M2\": %.17g, "
                "\"M1\": %.17g, \"wall_s\": %.6f}\n", e0, e2, e1, wall);
    return 0;
}
