// bench_batched_loop.cpp -- times the batched form against a loop of the
// plain per-instance calls (the way a reference user evaluates many
// instances: one launch_aggregation_rhs + launch_breakage_rhs per instance).
//
// Input: a binary file written by bench.py:
//   int32 count, int32 total, int32 n_quad,
//   count x {int32 n, int32 offset, 9 x float64 (log_x0, inv_log_r, beta0,
//            beta_br, beta_sh, S0, v_ref, alpha, v_min)},
//   float64 N[total], float64 x[total], float64 t_q[n_quad], float64 bw_q[n_quad]
// Output: one JSON line with the per-evaluation times and their ratio, and the
// largest |batched - loop| relative to max |loop|.
#include <pbe_cuda/pbe_cuda.cuh>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));   \
            std::exit(2);                                                          \
        }                                                                          \
    } while (0)

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s input.bin [reps]\n", argv[0]);
        return 2;
    }
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, f) != 3) return 2;
    const int count = hdr[0], total = hdr[1], nq = hdr[2];
    std::vector<pbe_batch_instance> inst(count);
    for (int i = 0; i < count; ++i) {
        int ni[2];
        double v[9];
        if (std::fread(ni, sizeof(int), 2, f) != 2 || std::fread(v, sizeof(double), 9, f) != 9)
            return 2;
        inst[i] = pbe_batch_instance{ni[0], ni[1], v[0], v[1], v[2], v[3], v[4], v[5], v[6],
                                     v[7], v[8]};
    }
    std::vector<double> N(total), x(total), tq(nq), bw(nq);
    if (std::fread(N.data(), sizeof(double), total, f) != (size_t)total ||
        std::fread(x.data(), sizeof(double), total, f) != (size_t)total ||
        std::fread(tq.data(), sizeof(double), nq, f) != (size_t)nq ||
        std::fread(bw.data(), sizeof(double), nq, f) != (size_t)nq)
        return 2;
    std::fclose(f);

    double *dN, *dx, *dr1, *dr2, *dt, *db;
    CHECK(hipMalloc(&dN, total * sizeof(double)));
    CHECK(hipMalloc(&dx, total * sizeof(double)));
    CHECK(hipMalloc(&dr1, total * sizeof(double)));
    CHECK(hipMalloc(&dr2, total * sizeof(double)));
    CHECK(hipMalloc(&dt, nq * sizeof(double)));
    CHECK(hipMalloc(&db, nq * sizeof(double)));
    CHECK(hipMemcpy(dN, N.data(), total * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dx, x.data(), total * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dt, tq.data(), nq * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(db, bw.data(), nq * sizeof(double), hipMemcpyHostToDevice));

    pbe_batch_plan plan;
    CHECK(pbe_batch_plan_create(&plan, inst.data(), count, PBE_AGG_BROWNIAN_SHEAR,
                                PBE_SEL_POWER_LAW, nq));
    hipStream_t s;
    CHECK(hipStreamCreate(&s));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));

    auto loop = [&]() {
        CHECK(hipMemsetAsync(dr2, 0, total * sizeof(double), s));
        for (int i = 0; i < count; ++i) {
            const pbe_batch_instance& b = inst[i];
            pbe_cuda::AggregationParams ap;
            ap.n = b.n;
            ap.log_x0 = b.log_x0;
            ap.inv_log_r = b.inv_log_r;
            ap.kernel_type = pbe_cuda::AggregationKernel::BrownianShear;
            ap.beta0 = b.beta0;
            ap.beta_br = b.beta_br;
            ap.beta_sh = b.beta_sh;
            pbe_cuda::BreakageParams bp;
            bp.n = b.n;
            bp.n_quad = nq;
            bp.selection = pbe_cuda::BreakageSelection::PowerLaw;
            bp.S0 = b.S0;
            bp.v_ref = b.v_ref;
            bp.alpha = b.alpha;
            bp.v_min = b.v_min;
            CHECK(pbe_cuda::launch_aggregation_rhs(dN + b.offset, dx + b.offset,
                                                   dr2 + b.offset, ap, s));
            CHECK(pbe_cuda::launch_breakage_rhs(dN + b.offset, dx + b.offset, dt, db,
                                                dr2 + b.offset, bp, s));
        }
    };
    auto batched = [&]() {
        CHECK(hipMemsetAsync(dr1, 0, total * sizeof(double), s));
        CHECK(pbe_launch_batched_rhs(plan, dN, dx, dt, db, dr1, s));
    };
    // warm up both, then time each
    loop();
    batched();
    CHECK(hipStreamSynchronize(s));
    float ms_loop = 0.f, ms_batch = 0.f;
    CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) loop();
    CHECK(hipEventRecord(e1, s));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms_loop, e0, e1));
    CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) batched();
    CHECK(hipEventRecord(e1, s));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms_batch, e0, e1));

    std::vector<double> r1(total), r2(total);
    CHECK(hipMemcpy(r1.data(), dr1, total * sizeof(double), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(r2.data(), dr2, total * sizeof(double), hipMemcpyDeviceToHost));
    double diff = 0.0, scale = 0.0;
    for (int i = 0; i < total; ++i) {
        diff = std::max(diff, std::fabs(r1[i] - r2[i]));
        scale = std::max(scale, std::fabs(r2[i]));
    }
    std::printf("{\"instances\": %d, \"bins\": %d, \"reps\": %d, \"ms_loop\": %.6f, "
                "\"ms_batched\": %.6f, \"ratio\": %.3f, \"max_rel_diff\": %.3e}\n",
                count, total, reps, ms_loop / reps, ms_batch / reps,
                (double)ms_loop / (double)ms_batch, scale > 0 ? diff / scale : 0.0);
    CHECK(pbe_batch_plan_destroy(plan));
    return 0;
}
