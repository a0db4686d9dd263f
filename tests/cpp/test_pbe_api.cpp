// test_pbe_api.cpp -- the reference's unit tests restated against the HIP
// drop-in C++ API (<pbe_cuda/pbe_cuda.cuh>).
//
// Mirrors tests/unit/test_aggregation.cpp (17 tests) and
// tests/unit/test_breakage.cpp (14 tests) of acuoci/pbe-fixed-pivot-cuda:
// same fixtures (n = 64, x_0 = 1e-18, R = 1.122018, geometric grid), same
// inputs, same assertions and tolerances.  GoogleTest is not available, so a
// minimal runner prints one line per test and exits non-zero on failure.
#include <pbe_cuda/pbe_cuda.cuh>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

namespace {

int g_failures = 0;
int g_checks_failed_in_test = 0;

#define EXPECT_TRUE_MSG(cond, ...)                                           \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ++g_checks_failed_in_test;                                       \
            std::printf("    check failed (%s:%d): %s  ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                                        \
            std::printf("\n");                                               \
        }                                                                    \
    } while (0)

#define HIP_OK(call)                                                         \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            ++g_checks_failed_in_test;                                       \
            std::printf("    HIP error %s at %s:%d\n", hipGetErrorString(e_),\
                        __FILE__, __LINE__);                                 \
        }                                                                    \
    } while (0)

struct Test {
    const char* name;
    std::function<void()> body;
};

std::vector<Test>& registry()
{
    static std::vector<Test> r;
    return r;
}

struct Reg {
    Reg(const char* n, std::function<void()> b) { registry().push_back({n, b}); }
};

#define TEST_CASE(name) \
    static void name(); \
    static Reg reg_##name(#name, name); \
    static void name()

// ---------------------------------------------------------------- fixture --
constexpr int    kN    = 64;
constexpr double kVmin = 1.0e-18;
constexpr double kR    = 1.122018;

struct Grid {
    std::vector<double> x, N, rhs;
    double *d_x = nullptr, *d_N = nullptr, *d_rhs = nullptr;
    double *d_tq = nullptr, *d_bw = nullptr;
    double log_x0 = 0.0, inv_log_r = 0.0;
    std::vector<double> tq{0.5}, bw{2.0};

    Grid() : x(kN), N(kN, 0.0), rhs(kN, 0.0)
    {
        x[0] = kVmin;
        for (int i = 1; i < kN; ++i) x[i] = x[i - 1] * kR;
        log_x0 = std::log(x[0]);
        inv_log_r = 1.0 / std::log(x[1] / x[0]);
        HIP_OK(hipMalloc(&d_x, kN * sizeof(double)));
        HIP_OK(hipMalloc(&d_N, kN * sizeof(double)));
        HIP_OK(hipMalloc(&d_rhs, kN * sizeof(double)));
        HIP_OK(hipMalloc(&d_tq, sizeof(double)));
        HIP_OK(hipMalloc(&d_bw, sizeof(double)));
        HIP_OK(hipMemcpy(d_x, x.data(), kN * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_tq, tq.data(), sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_bw, bw.data(), sizeof(double), hipMemcpyHostToDevice));
    }
    ~Grid()
    {
        hipFree(d_x); hipFree(d_N); hipFree(d_rhs); hipFree(d_tq); hipFree(d_bw);
    }

    pbe_cuda::AggregationParams agg(pbe_cuda::AggregationKernel k, double b0 = 1.0e-17,
                                    double b_br = 0.0, double b_sh = 0.0) const
    {
        pbe_cuda::AggregationParams p;
        p.n = kN; p.log_x0 = log_x0; p.inv_log_r = inv_log_r; p.kernel_type = k;
        p.beta0 = b0; p.beta_br = b_br; p.beta_sh = b_sh; p.block_size = 256;
        return p;
    }
    pbe_cuda::BreakageParams brk(pbe_cuda::BreakageSelection s, double S0 = 1.0e-3,
                                 double v_ref = 1.0e-18, double alpha = 1.0,
                                 double v_min = 0.0) const
    {
        pbe_cuda::BreakageParams p;
        p.n = kN; p.n_quad = 1; p.selection = s; p.S0 = S0; p.v_ref = v_ref;
        p.alpha = alpha; p.v_min = v_min; p.block_size = 256;
        return p;
    }
    hipError_t run(const pbe_cuda::AggregationParams& p)
    {
        HIP_OK(hipMemcpy(d_N, N.data(), kN * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemset(d_rhs, 0, kN * sizeof(double)));
        hipError_t e = pbe_cuda::launch_aggregation_rhs(d_N, d_x, d_rhs, p);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(rhs.data(), d_rhs, kN * sizeof(double), hipMemcpyDeviceToHost));
        return e;
    }
    hipError_t run(const pbe_cuda::BreakageParams& p)
    {
        HIP_OK(hipMemcpy(d_N, N.data(), kN * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemset(d_rhs, 0, kN * sizeof(double)));
        hipError_t e = pbe_cuda::launch_breakage_rhs(d_N, d_x, d_tq, d_bw, d_rhs, p);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(rhs.data(), d_rhs, kN * sizeof(double), hipMemcpyDeviceToHost));
        return e;
    }
    double dM1() const
    {
        double s = 0.0;
        for (int i = 0; i < kN; ++i) s += rhs[i] * x[i];
        return s;
    }
    double M1() const
    {
        double s = 0.0;
        for (int i = 0; i < kN; ++i) s += N[i] * x[i];
        return s;
    }
};

using AK = pbe_cuda::AggregationKernel;
using BS = pbe_cuda::BreakageSelection;

// ===================== aggregation (tests/unit/test_aggregation.cpp) ======
void agg_smoke(AK k, double b0, double bbr, double bsh)
{
    Grid g;
    g.N[kN / 2] = 1.0e14;
    EXPECT_TRUE_MSG(g.run(g.agg(k, b0, bbr, bsh)) == hipSuccess, "launch failed");
}
TEST_CASE(AggregationTest_SmokeConstant) { agg_smoke(AK::Constant, 1e-17, 0, 0); }
TEST_CASE(AggregationTest_SmokeSum) { agg_smoke(AK::Sum, 1e-17, 0, 0); }
TEST_CASE(AggregationTest_SmokeProduct) { agg_smoke(AK::Product, 1e-17, 0, 0); }
TEST_CASE(AggregationTest_SmokeBrownian) { agg_smoke(AK::Brownian, 0.0, 1e-17, 0.0); }
TEST_CASE(AggregationTest_SmokeShear) { agg_smoke(AK::Shear, 0.0, 0.0, 1e-17); }
TEST_CASE(AggregationTest_SmokeBrownianShear) { agg_smoke(AK::BrownianShear, 0.0, 1e-17, 1e-17); }

void agg_conservation(AK k, double b0, double bbr, double bsh)
{
    Grid g;
    for (int i = 0; i < kN; ++i) g.N[i] = 1.0e12 * std::exp(-static_cast<double>(i) / 10.0);
    EXPECT_TRUE_MSG(g.run(g.agg(k, b0, bbr, bsh)) == hipSuccess, "launch failed");
    const double tol = 1.0e-6 * std::abs(g.M1());
    EXPECT_TRUE_MSG(std::abs(g.dM1()) <= tol, "dM1 = %g, M1 = %g", g.dM1(), g.M1());
}
TEST_CASE(AggregationConservationTest_VolumeConservedConstant) { agg_conservation(AK::Constant, 1e-17, 0, 0); }
TEST_CASE(AggregationConservationTest_VolumeConservedSum) { agg_conservation(AK::Sum, 1e-17, 0, 0); }
TEST_CASE(AggregationConservationTest_VolumeConservedProduct) { agg_conservation(AK::Product, 1e-17, 0, 0); }
TEST_CASE(AggregationConservationTest_VolumeConservedBrownian) { agg_conservation(AK::Brownian, 0.0, 1e-17, 0.0); }
TEST_CASE(AggregationConservationTest_VolumeConservedShear) { agg_conservation(AK::Shear, 0.0, 0.0, 1e-17); }
TEST_CASE(AggregationConservationTest_VolumeConservedBrownianShear) { agg_conservation(AK::BrownianShear, 0.0, 1e-17, 1e-17); }

TEST_CASE(AggregationTest_ZeroPopulationGivesZeroRHS)
{
    Grid g;
    EXPECT_TRUE_MSG(g.run(g.agg(AK::Constant)) == hipSuccess, "launch failed");
    for (int i = 0; i < kN; ++i) EXPECT_TRUE_MSG(g.rhs[i] == 0.0, "rhs[%d] = %g", i, g.rhs[i]);
}

TEST_CASE(AggregationTest_SelfCollisionDeathConstantKernel)
{
    Grid g;
    const int j = kN / 2;
    const double Nj = 1.0e14, b0 = 1.0e-17;
    g.N[j] = Nj;
    EXPECT_TRUE_MSG(g.run(g.agg(AK::Constant, b0)) == hipSuccess, "launch failed");
    const double expected = -b0 * Nj * Nj;
    EXPECT_TRUE_MSG(std::abs(g.rhs[j] - expected) <= std::abs(expected) * 1.0e-10,
                    "rhs[%d] = %.17g expected %.17g", j, g.rhs[j], expected);
}

TEST_CASE(AggregationTest_TotalNumberNonIncreasing)
{
    Grid g;
    for (int i = 0; i < kN; ++i) g.N[i] = 1.0e12;
    EXPECT_TRUE_MSG(g.run(g.agg(AK::Constant)) == hipSuccess, "launch failed");
    const double s = std::accumulate(g.rhs.begin(), g.rhs.end(), 0.0);
    EXPECT_TRUE_MSG(s <= 0.0, "sum rhs = %g", s);
}

TEST_CASE(AggregationTest_InvalidNReturnsError)
{
    Grid g;
    auto p = g.agg(AK::Constant);
    p.n = 0;
    EXPECT_TRUE_MSG(pbe_cuda::launch_aggregation_rhs(g.d_N, g.d_x, g.d_rhs, p) != hipSuccess,
                    "n = 0 accepted");
}

TEST_CASE(AggregationTest_NullPointerReturnsError)
{
    Grid g;
    auto p = g.agg(AK::Constant);
    EXPECT_TRUE_MSG(pbe_cuda::launch_aggregation_rhs(nullptr, g.d_x, g.d_rhs, p) != hipSuccess,
                    "null N accepted");
}

// ========================= breakage (tests/unit/test_breakage.cpp) ========
void brk_smoke(BS s, double S0, double v_ref, double alpha, double v_min)
{
    Grid g;
    g.N[kN / 2] = 1.0e14;
    EXPECT_TRUE_MSG(g.run(g.brk(s, S0, v_ref, alpha, v_min)) == hipSuccess, "launch failed");
}
TEST_CASE(BreakageTest_SmokeConstant) { brk_smoke(BS::Constant, 1e-3, 1e-18, 1.0, 0.0); }
TEST_CASE(BreakageTest_SmokeLinear) { Grid g; brk_smoke(BS::Linear, 1e-3, g.x[kN / 2], 1.0, 0.0); }
TEST_CASE(BreakageTest_SmokePowerLaw) { Grid g; brk_smoke(BS::PowerLaw, 1e-3, g.x[kN / 2], 2.0, 0.0); }
TEST_CASE(BreakageTest_SmokeThreshold) { Grid g; brk_smoke(BS::Threshold, 1e-3, 1.0, 1.0, g.x[kN / 4]); }

void brk_conservation(BS s, double S0, double v_ref, double alpha, double v_min)
{
    Grid g;
    for (int i = 0; i < kN; ++i) g.N[i] = 1.0e12 * std::exp(-static_cast<double>(i) / 10.0);
    EXPECT_TRUE_MSG(g.run(g.brk(s, S0, v_ref, alpha, v_min)) == hipSuccess, "launch failed");
    const double tol = 1.0e-4 * std::abs(g.M1());
    EXPECT_TRUE_MSG(std::abs(g.dM1()) <= tol, "dM1 = %g, M1 = %g", g.dM1(), g.M1());
}
TEST_CASE(BreakageConservationTest_VolumeConservedConstant) { brk_conservation(BS::Constant, 1e-3, 1e-18, 1.0, 0.0); }
TEST_CASE(BreakageConservationTest_VolumeConservedLinear) { Grid g; brk_conservation(BS::Linear, 1e-3, g.x[kN / 2], 1.0, 0.0); }
TEST_CASE(BreakageConservationTest_VolumeConservedPowerLaw) { Grid g; brk_conservation(BS::PowerLaw, 1e-3, g.x[kN / 2], 2.0, 0.0); }
TEST_CASE(BreakageConservationTest_VolumeConservedThreshold) { Grid g; brk_conservation(BS::Threshold, 1e-3, 1.0, 1.0, g.x[kN / 4]); }

TEST_CASE(BreakageTest_ZeroPopulationGivesZeroRHS)
{
    Grid g;
    EXPECT_TRUE_MSG(g.run(g.brk(BS::Constant)) == hipSuccess, "launch failed");
    for (int i = 0; i < kN; ++i) EXPECT_TRUE_MSG(g.rhs[i] == 0.0, "rhs[%d] = %g", i, g.rhs[i]);
}

TEST_CASE(BreakageTest_DeathTermConstantSelection)
{
    Grid g;
    const int j = kN / 2;
    const double Nj = 1.0e14, S0 = 1.0e-3;
    g.N[j] = Nj;
    EXPECT_TRUE_MSG(g.run(g.brk(BS::Constant, S0)) == hipSuccess, "launch failed");
    const double expected = -S0 * Nj;
    EXPECT_TRUE_MSG(std::abs(g.rhs[j] - expected) <= std::abs(expected) * 1.0e-10,
                    "rhs[%d] = %.17g expected %.17g", j, g.rhs[j], expected);
}

TEST_CASE(BreakageTest_TotalNumberNonDecreasing)
{
    Grid g;
    for (int i = 0; i < kN; ++i) g.N[i] = 1.0e12;
    EXPECT_TRUE_MSG(g.run(g.brk(BS::Constant)) == hipSuccess, "launch failed");
    const double s = std::accumulate(g.rhs.begin(), g.rhs.end(), 0.0);
    EXPECT_TRUE_MSG(s >= 0.0, "sum rhs = %g", s);
}

TEST_CASE(BreakageTest_ThresholdSuppressesBelowVmin)
{
    Grid g;
    const int tb = kN / 2;
    for (int i = 0; i < kN; ++i) g.N[i] = 1.0e12;
    EXPECT_TRUE_MSG(g.run(g.brk(BS::Threshold, 1.0e-3, 1.0, 1.0, g.x[tb])) == hipSuccess,
                    "launch failed");
    for (int i = 0; i < tb; ++i) EXPECT_TRUE_MSG(g.rhs[i] >= 0.0, "rhs[%d] = %g", i, g.rhs[i]);
}

TEST_CASE(BreakageTest_InvalidNReturnsError)
{
    Grid g;
    auto p = g.brk(BS::Constant);
    p.n = 0;
    EXPECT_TRUE_MSG(pbe_cuda::launch_breakage_rhs(g.d_N, g.d_x, g.d_tq, g.d_bw, g.d_rhs, p)
                    != hipSuccess, "n = 0 accepted");
}

TEST_CASE(BreakageTest_NullPointerReturnsError)
{
    Grid g;
    auto p = g.brk(BS::Constant);
    EXPECT_TRUE_MSG(pbe_cuda::launch_breakage_rhs(nullptr, g.d_x, g.d_tq, g.d_bw, g.d_rhs, p)
                    != hipSuccess, "null N accepted");
}

}  // namespace

int main(int argc, char** argv)
{
    const char* filter = argc > 1 ? argv[1] : nullptr;
    int run = 0;
    for (auto& t : registry()) {
        if (filter && std::string(t.name).find(filter) == std::string::npos) continue;
        g_checks_failed_in_test = 0;
        t.body();
        ++run;
        std::printf("[%s] %s\n", g_checks_failed_in_test ? "FAILED" : "    OK", t.name);
        std::fflush(stdout);
        if (g_checks_failed_in_test) ++g_failures;
    }
    std::printf("%d tests, %d failed\n", run, g_failures);
    return g_failures ? 1 : 0;
}
