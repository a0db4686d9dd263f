// test_roots.cpp -- host check of cbrt_pair (pbe_device.hpp) against the C
// library's cbrt: worst error in ulp over log-uniform x in [1e-300, 1e300]
// plus edge values.  Runs without a GPU.
#include "pbe_device.hpp"

#include <cmath>
#include <cstdio>
#include <random>

static double ulps(double got, double ref)
{
    const double u = std::nextafter(ref, INFINITY) - ref;
    return std::fabs(got - ref) / u;
}

int main()
{
    std::mt19937_64 rng(20261203);
    std::uniform_real_distribution<double> e(-300.0, 300.0), m(0.0, 1.0);
    double wc = 0.0, wi = 0.0;
    for (int i = 0; i < 4000000; ++i) {
        const double x = (1.0 + m(rng)) * std::pow(10.0, e(rng));
        double c, ic;
        pbe::cbrt_pair(x, c, ic);
        const double rc = std::cbrt(x);
        wc = std::fmax(wc, ulps(c, rc));
        wi = std::fmax(wi, ulps(ic, 1.0 / rc));
    }
    const double edge[] = {1e-18, 1.0, 8.0, 27.0, 2.0, 1e-300, 1e300, 2.2250738585072014e-308,
                           1.7976931348623157e308, 1e-310, 0.0};
    for (double x : edge) {
        double c, ic;
        pbe::cbrt_pair(x, c, ic);
        const double rc = std::cbrt(x);
        if (x > 0.0) {
            wc = std::fmax(wc, ulps(c, rc));
            wi = std::fmax(wi, ulps(ic, 1.0 / rc));
        } else if (c != 0.0) {
            wc = 1e9;
        }
    }
    std::printf("{\"cbrt_max_ulp\": %.2f, \"inv_cbrt_max_ulp\": %.2f}\n", wc, wi);
    return (wc <= 4.0 && wi <= 8.0) ? 0 : 1;
}
