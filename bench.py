#!/usr/bin/env python3
"""Benchmark: FP64 pair-updates/s of one aggregation + breakage evaluation.

Workload (one step): rhs zeroed on the stream, then pbe_launch_aggregation_rhs
(Brownian+shear kernel) and pbe_launch_breakage_rhs (threshold selection,
symmetric-binary daughters) at n = 16384 on a geometric grid r = 1.02, with a
seeded lognormal N already resident in HBM.  Work units per step:
n(n+1)/2 pairs + n*n_quad breakage updates (dense: pairs with N = 0 count).

Prints one JSON line.  ``roofline`` refers to the dominant work, the
aggregation call (at n = 16384 one fused launch of the diagonal-band sweep
items and the column strips): achieved = 21 FP64 FLOP per pair whose birth
weight is not exactly zero (fl(x_j + x_k) != x_k; the rest only lose their
rate from bin j, a separable O(n) death) / average duration of the call, K calls back to back between two HIP events on
the launch stream (the figure rocprofv3's kernel trace of the launch agrees
with); beside it the average of HIP events around each call in a second pass
over the same steps (events between launches add their own cost, so the
timed region carries events only at its two ends), and the FP64 FLOPs the
call executes (rocprofv3
SQ_INSTS_VALU_FLOPS_FP64 of the same launch, committed under profiles/) over
the same time.
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
TRAFFIC_FILE = os.path.join(REPO, "profiles", "agg_brshear_n16384_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=2.0,
                    help="wall-clock budget of the CPU baseline sample")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched and adjoint comparisons")
    ap.add_argument("--no-fit", action="store_true",
                    help="skip the batched joint fit (its thousands of launches swamp a trace)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the n = 1024 / 4096 / GL16 / 65536 variants (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run an elementwise device copy of BYTES (torch.neg, "
                         "reads and writes BYTES each), for scaling HBM byte counters")
    return ap.parse_args()


def copy_calibration(nbytes):
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": src.numel() * 8, "launches": 3,
                      "kernel": "torch.neg (elementwise)"}))


def batched_form(hip, P):
    """4096 instances (n log-uniform in [64, 2048]): one batched launch vs a
    loop of plain per-instance calls, timed by the C++ helper on one stream."""
    import struct
    import subprocess
    import tempfile
    inst, N, x, t_q, bw_q = P.config_batched(4096, seed=4096)
    exe = os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "bench", "batched_loop")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "batch.bin")
        with open(path, "wb") as f:
            f.write(struct.pack("iii", len(inst), len(N), len(t_q)))
            for d in inst:
                f.write(struct.pack("ii9d", d["n"], d["offset"], d["log_x0"], d["inv_log_r"],
                                    d["beta0"], d["beta_br"], d["beta_sh"], d["S0"], d["v_ref"],
                                    d["alpha"], d["v_min"]))
            for a in (N, x, t_q, bw_q):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([exe, path, "3"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=240)
        if r.returncode != 0:
            return {"error": r.stderr[-500:]}
        out = json.loads(r.stdout.strip().splitlines()[-1])
    pairs = sum(d["n"] * (d["n"] + 1) // 2 for d in inst)
    out["pairs_per_eval"] = pairs
    out["batched_pair_updates_per_s"] = (pairs + len(N)) / (out["ms_batched"] * 1e-3)
    out["workload"] = ("4096 instances, n log-uniform [64, 2048], r = 1.03, own G in [10, 500], "
                       "Brownian+shear + power-law breakage, binary daughters")
    return out


def adjoint_form(pbe, hip, P):
    """n = 1024, 2000 RK4 steps of 0.5 s: gradient (adjoint forward + reverse)
    vs a plain forward solve, all timed with HIP events after one untimed
    warm-up of each.  Two adjoint modes: checkpoints every 50 steps with
    segment recompute, and every stage state kept in HBM (4 x 2000 x 1024
    doubles = 66 MB) with no recompute."""
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(1024)
    dt, steps = 0.5, 2000
    s = hip.Stream()
    it = pbe.Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, dt, 5, s))
    s.synchronize()
    dN.upload(N0)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    hip.check(it.step(dN, dt, steps, s))
    e1.record(s)
    s.synchronize()
    t_fwd = e0.elapsed_ms(e1)
    out = {"n": 1024, "steps": steps, "forward_solve_ms": t_fwd,
           "workload": "Brownian+shear aggregation + power-law breakage, r = 1.03, n = 1024"}
    best = None
    for every, name in ((50, "checkpoint_50"), (pbe.STORE_ALL_STAGES, "store_all_stages")):
        adj = pbe.Adjoint(x, dt, steps, every, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
        dA = hip.DeviceArray.from_host(N0)
        hip.check(adj.forward(dA, s))
        s.synchronize()
        Nf = dA.download()
        logT = np.log(np.maximum(Nf * 1.1, 1e-26))
        mask = (Nf > 1e-8 * Nf.max()).astype(float)
        L, gfin = P.log_misfit(Nf, logT, mask, 1e-26)
        dg, gN0 = hip.DeviceArray.from_host(gfin), hip.DeviceArray(len(N0))
        gth = hip.DeviceArray(5)
        hip.check(adj.backward(dg, gN0, gth, s))   # warm-up pass
        s.synchronize()
        dA.upload(N0)
        g0, g1, g2, h0 = hip.Event(), hip.Event(), hip.Event(), hip.Event()
        g0.record(s)
        hip.check(adj.forward(dA, s))
        g1.record(s)
        s.synchronize()
        h0.record(s)
        hip.check(adj.backward(dg, gN0, gth, s))
        g2.record(s)
        s.synchronize()
        t_afwd, t_abwd = g0.elapsed_ms(g1), h0.elapsed_ms(g2)
        ratio = (t_afwd + t_abwd) / t_fwd
        out[name] = {"adjoint_forward_ms": t_afwd, "adjoint_backward_ms": t_abwd,
                     "workspace_MB": adj.ws_bytes / 1e6, "gradient_over_forward": ratio}
        if best is None or ratio < best[0]:
            best = (ratio, name)
    out["gradient_over_forward"], out["best_mode"] = best
    return out


def batched_fit_form(hip, P):
    """The adjoint workload's batched variant: 256 instances, n log-uniform
    in [128, 1024], each with its own G and initial PSD, fitted jointly.
    Times one forward solve and one gradient (batched adjoint forward +
    reverse; checkpoints every 50 steps with recompute, or every stage state
    kept in HBM: 4 x 2000 x 114514 doubles, 7.3 GB) at 2000 RK4 steps of
    0.5 s, each after an untimed warm-up, and runs an
    L-BFGS fit of the common factors at 200 steps from all factors 1 to a
    target solved at known ones."""
    import pbe_hip
    from pbe_hip.fit import BatchedFit
    inst, N0, x, t_q, bw_q = P.config_batched_fit(256, seed=256)
    rng = np.random.default_rng(11)
    th_true = np.array([*np.log(rng.uniform(0.5, 2.0, 3)), (1.0 / 3.0) * rng.uniform(0.5, 2.0)])
    out = {"instances": len(inst), "bins": len(N0),
           "pairs_per_eval": sum(d["n"] * (d["n"] + 1) // 2 for d in inst),
           "workload": ("256 instances, n log-uniform [128, 1024], r = 1.03, own G in [0.5, 2], "
                        "Brownian+shear + power-law breakage, binary daughters; fitted: common "
                        "factors of beta_br, beta_sh, S0 and the exponent alpha")}
    F = BatchedFit(inst, N0, x, t_q, bw_q, 0.5, 2000, every=50)
    F.set_target(th_true)          # warm-up of both passes
    F.loss_and_grad(F.theta0)
    t0 = time.perf_counter()
    F.set_target(th_true)
    t1 = time.perf_counter()
    out["steps"] = 2000
    out["forward_solve_s"] = t1 - t0
    best = None
    for every, name in ((50, "checkpoint_50"), (0, "store_all_stages")):
        Fm = F
        if every != 50:
            Fm = BatchedFit(inst, N0, x, t_q, bw_q, 0.5, 2000, every=every)
            Fm.logT, Fm.mask = F.logT, F.mask
            Fm.loss_and_grad(Fm.theta0)   # warm-up
        t2 = time.perf_counter()
        Fm.loss_and_grad(Fm.theta0)
        t3 = time.perf_counter()
        ratio = (t3 - t2) / (t1 - t0)
        out[name] = {"gradient_s": t3 - t2, "gradient_over_forward": ratio,
                     "workspace_GB": pbe_hip.library().pbe_adjoint_workspace_bytes(
                         len(N0), 2000, every) / 1e9}
        if best is None or ratio < best[0]:
            best = (ratio, name)
    out["gradient_over_forward"], out["best_mode"] = best
    G = BatchedFit(inst, N0, x, t_q, bw_q, 0.5, 200, every=50)
    G.set_target(th_true)
    t3 = time.perf_counter()
    th, hist, res = G.fit(max_evals=60)
    t4 = time.perf_counter()
    out["fit"] = {"steps": 200, "evaluations": len(hist), "seconds": t4 - t3,
                  "loss_start": hist[0], "loss_end": hist[-1],
                  "max_abs_theta_error": float(np.max(np.abs(th - th_true)))}
    return out


def rhs_variant(pbe_hip, hip, P, n, daughter, steps=20, r=1.02):
    """Average step time (zero rhs, aggregation, breakage) of one flocculation
    RHS configuration, HIP events on the launch stream."""
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(n, daughter=daughter, n_quad=16, r=r)
    nq = len(t_q)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
    drhs = hip.DeviceArray(n)
    ap, bp = pbe_hip.agg_params(n, agg), pbe_hip.brk_params(n, nq, brk)
    s = hip.Stream()
    rt = hip.runtime()

    def step():
        hip.check(rt.hipMemsetAsync(drhs.ptr, 0, drhs.nbytes, s.handle), "memset")
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, s), "agg")
        hip.check(pbe_hip.launch_breakage_rhs(dN, dx, dt, db, drhs, bp, s), "brk")

    for _ in range(3):
        step()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    for _ in range(steps):
        step()
    e1.record(s)
    s.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    return {"n": n, "r": r, "daughter": daughter, "n_quad": nq, "ms_per_step": ms,
            "pair_updates_per_s": P.pair_units(n, nq) / (ms * 1e-3)}


def deterministic_form(pbe_hip, hip, P, n=16384, steps=10):
    """The atomic-free gather launches on the default workload: average step
    time (zero rhs, aggregation, breakage) against the atomic launches."""
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(n)
    nq = len(t_q)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
    drhs = hip.DeviceArray(n)
    ap, bp = pbe_hip.agg_params(n, agg), pbe_hip.brk_params(n, nq, brk)
    ws_bytes = max(pbe_hip.aggregation_deterministic_workspace_bytes(n),
                   pbe_hip.breakage_deterministic_workspace_bytes(n))
    ws = hip.DeviceArray(ws_bytes // 8 + 1)
    s = hip.Stream()
    rt = hip.runtime()

    def step():
        hip.check(rt.hipMemsetAsync(drhs.ptr, 0, drhs.nbytes, s.handle), "memset")
        hip.check(pbe_hip.launch_aggregation_rhs_deterministic(dN, dx, drhs, ap, ws, s), "agg")
        hip.check(pbe_hip.launch_breakage_rhs_deterministic(dN, dx, dt, db, drhs, bp, ws, s),
                  "brk")

    for _ in range(2):
        step()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    for _ in range(steps):
        step()
    e1.record(s)
    s.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    return {"n": n, "ms_per_step": ms, "workspace_bytes": ws_bytes,
            "pair_updates_per_s": P.pair_units(n, nq) / (ms * 1e-3)}


def flocculation_form():
    """The reference's two-stage flocculation case study end to end
    (examples/flocculation_two_stage, 3600 RK4 steps, G 100 -> 20 1/s, the
    time loop on the device), at N = 960 and N = 3840: wall time of the
    program's loop (its 60 prints with host copies included, as in the
    reference's) beside the reference's published A30 figures."""
    import subprocess
    exe = os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "examples", "flocculation_two_stage",
                       "example_flocculation_two_stage")
    published = {960: ("~4 s on an NVIDIA A30 (examples/flocculation_two_stage/main.cpp:54)", 4.0),
                 3840: ("37.6 s on an NVIDIA A30 (README.md:43)", 37.6)}
    out = []
    for n in (960, 3840):
        r = subprocess.run([exe, str(n)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=240)
        if r.returncode != 0:
            out.append({"n": n, "error": r.stderr[-300:]})
            continue
        d = json.loads(r.stdout.strip().splitlines()[-1])
        d["reference_published"], ref_s = published[n]
        d["speedup_vs_published"] = ref_s / d["wall_s"]
        out.append(d)
    return out


def spot_check(got, N, x, agg, brk, t_q, bw_q):
    """Every entry of one GPU step against the OpenMP port, within
    1e-10 max|port| + 1e-9 |port_i| (the port sums in plain FP64 in another
    order), and the first moment within 1e-13 of sum x |port|."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    port = oracle.port_aggregation_rhs(N, x, threads=threads, **agg)
    port += oracle.port_breakage_rhs(N, x, t_q, bw_q, threads=threads, **brk)
    err = np.abs(got - port) / (1e-10 * np.abs(port).max() + 1e-9 * np.abs(port))
    m1 = abs(float(np.sum(x * got)) - float(np.sum(x * port))) / float(np.sum(x * np.abs(port)))
    if not (err.max() <= 1.0 and m1 <= 1e-13):
        raise SystemExit(f"bench spot check failed: entry {err.max():.3e} (of the bound), "
                         f"first moment {m1:.3e}")
    return {"vs": "OpenMP port (oracle/pbe_cpu_port.c)", "max_entry_over_bound": float(err.max()),
            "first_moment_rel": m1}


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
    # the same on one core: one evaluation
    t1 = time.perf_counter()
    r = oracle.port_aggregation_rhs(N, x, threads=1, **agg)
    r += oracle.port_breakage_rhs(N, x, t_q, bw_q, threads=1, **brk)
    el1 = time.perf_counter() - t1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f
                          if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": units * evals / el,
        "unit": "pair-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{evals} full agg+breakage RHS evaluation(s) at n={n} "
                   f"({el:.2f} s wall, {threads} OpenMP threads, "
                   "C restatement of the reference pair loop)"),
        "one_core_value": units / el1,
        "one_core_sample": f"1 evaluation on 1 thread ({el1:.2f} s wall)",
        "cpu_model": model,
        "host_cpus_visible": os.cpu_count(),
    }


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                          "(the path is not sharded across GPUs here)",
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

    # timed region: K steps back to back on the launch stream, events only
    # at its two ends (an event recorded between two launches adds several
    # microseconds of its own to every step)
    K = args.steps
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(K):
        step()
    e1.record(stream)
    stream.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_ms(e1) / K

    # the aggregation launch's own duration: the same K steps again, with
    # events on the launch stream around each aggregation call
    ev = [(hip.Event(), hip.Event()) for _ in range(K)]
    g0, g1 = hip.Event(), hip.Event()
    g0.record(stream)
    for i in range(K):
        hip.check(rt.hipMemsetAsync(drhs.ptr, 0, drhs.nbytes, stream.handle), "memset")
        ev[i][0].record(stream)
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, stream), "agg")
        ev[i][1].record(stream)
        hip.check(pbe_hip.launch_breakage_rhs(dN, dx, dt, db, drhs, bp, stream), "brk")
    g1.record(stream)
    stream.synchronize()
    ms_step_events = g0.elapsed_ms(g1) / K
    agg_each = sorted(a.elapsed_ms(b) for a, b in ev)
    agg_ms = sum(agg_each) / K
    agg_median = agg_each[K // 2] if K % 2 else 0.5 * (agg_each[K // 2 - 1] + agg_each[K // 2])

    # K aggregation launches back to back between two events (rhs
    # accumulates; timing only): the launch's duration without the events'
    # own cost
    b0, b1 = hip.Event(), hip.Event()
    b0.record(stream)
    for _ in range(K):
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, stream), "agg")
    b1.record(stream)
    stream.synchronize()
    agg_ms_b2b = b0.elapsed_ms(b1) / K

    # same K steps without the per-step memset (rhs accumulates; timing only)
    f0, f1 = hip.Event(), hip.Event()
    f0.record(stream)
    for i in range(K):
        hip.check(pbe_hip.launch_aggregation_rhs(dN, dx, drhs, ap, stream), "agg")
        hip.check(pbe_hip.launch_breakage_rhs(dN, dx, dt, db, drhs, bp, stream), "brk")
    f1.record(stream)
    stream.synchronize()
    ms_step_nomemset = f0.elapsed_ms(f1) / K

    # correctness spot check of one evaluated step against the OpenMP port
    # (plain FP64, the reference's bracket rule), outside the timed region;
    # a mismatch fails the run
    drhs.zero()
    step()
    stream.synchronize()
    spot = spot_check(drhs.download(), N, x, agg, brk, t_q, bw_q)

    # algorithmic FLOPs: 21 per pair whose birth weight is not exactly zero
    # (fl(x_j + x_k) != x_k); the other pairs only move their rate out of
    # bin j, a separable death (O(n) for all of them), and are left out
    bpairs = P.birth_pairs(x)
    flops = P.PAIR_FLOPS[int(agg["kernel"])] * bpairs
    dense_flops = P.PAIR_FLOPS[int(agg["kernel"])] * pairs
    # the launch's average duration: K launches back to back between two
    # events (agrees with rocprofv3's kernel trace of the same launch; events
    # around each call add their own cost to every reading)
    achieved = flops / (agg_ms_b2b * 1e-3) / 1e12
    # the launches the library's tuned shape makes for this call
    tu, tw, tf = pbe_hip.aggregation_default_options(ap)
    kname = pbe_hip.AggregationKernel(int(agg["kernel"])).name
    if tu == 4 and tf in (2, 4) and -(-n // 256) <= 64:
        agg_label = (f"aggregation call: agg_fused_kernel<{kname}> (band split {tw}, "
                     f"{3 if tf == 4 else 2} workgroups per CU)")
    elif tf == 3:
        agg_label = f"aggregation call: agg_tiles_kernel<{kname}, {tu}> (band split {tw})"
    elif tu == 4:
        agg_label = (f"aggregation call: agg_sweep_kernel<{kname}> (band split {tw}) + "
                     f"agg_strip_kernel<{kname}, 4>")
    else:
        agg_label = (f"aggregation call: {64 * tu}-column tiles and strips <{kname}> "
                     f"(band split {tw})")
    # HBM bytes per call from the committed counter runs of the same launches
    # (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-kernel calibrated); null when
    # they were taken on other launches
    traffic = None
    executed = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tj = json.load(f)
        if tj.get("launch") == agg_label:
            traffic = tj.get("hbm_bytes_per_launch")
            executed = tj.get("fp64_flop_executed_per_launch")

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
            # compute-bound on the FP64 vector ALU (no matrix instructions:
            # not a contraction); MI355X's FP64 vector peak is 78.6 TFLOP/s
            "bound": "valu",
            "pipe": "fp64 VALU",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": agg_label,
            "flop_per_launch": flops,
            "birth_pairs": bpairs,
            "dense_flop_per_launch": dense_flops,
            "dense_tflops": dense_flops / (agg_ms_b2b * 1e-3) / 1e12,
            "executed_fp64_flop_per_launch": executed,
            "executed_tflops": (executed / (agg_ms_b2b * 1e-3) / 1e12) if executed else None,
            "executed_frac": (executed / (agg_ms_b2b * 1e-3) / 1e12 / FP64_PEAK_TFLOPS)
            if executed else None,
            "avg_launch_ms": agg_ms_b2b,
            "avg_launch_ms_events_each": agg_ms,
            "median_launch_ms_events_each": agg_median,
            "frac_events_each": flops / (agg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        },
        "ms_per_step_with_launch_events": ms_step_events,
        "ms_per_step_without_memset": ms_step_nomemset,
        "spot_check": spot,
        "wall_s_timed_region": wall,
    }
    if not args.no_variants:
        out["variants"] = [rhs_variant(pbe_hip, hip, P, m, "symmetric") for m in (1024, 4096)]
        out["variants"].append(rhs_variant(pbe_hip, hip, P, n, "uniform"))
        # a finer grid: near-pair reach 1062 bins
        out["variants"].append(rhs_variant(pbe_hip, hip, P, n, "symmetric", r=1.005))
        # past the int32 pair count (2.1e9 pairs); r = 1.005 keeps x finite
        out["variants"].append(rhs_variant(pbe_hip, hip, P, 65536, "symmetric", steps=5,
                                           r=1.005))
    if not args.no_forms:
        out["forms"] = {"batched": batched_form(hip, P), "adjoint": adjoint_form(pbe_hip, hip, P),
                        "deterministic": deterministic_form(pbe_hip, hip, P, n)}
        out["forms"]["deterministic"]["over_atomic"] = (
            out["forms"]["deterministic"]["ms_per_step"] / ms_step)
        out["forms"]["flocculation_two_stage"] = flocculation_form()
        if not args.no_fit:
            out["forms"]["batched_fit"] = batched_fit_form(hip, P)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, x, agg, brk, t_q, bw_q, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
