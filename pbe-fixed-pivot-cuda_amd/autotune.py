"""Pick the aggregation launch shapes by measurement and write the tables the
library compiles in.

    python pbe-fixed-pivot-cuda_amd/autotune.py [--part single|batch|both]
        [--out-dir DIR] [--summary FILE] [--quick]

Single-instance launches (csrc/tile_table.inc, rows {n_max, kernel_type,
reach_max, U, S, F}): for every collision kernel, grid-reach class and size
bucket, every block width 64*U (U = 1, 2, 4, 8 columns per lane) and band
split S = 1, 2, 4, 8, 16 (row slices per diagonal-band tile; offset segments
per block for the 256-wide band sweeps) is timed as two launches (band,
then strips; F = 1) and as one launch of every tile (F = 3), at U = 4 also
as the one fused band + strip launch at two or three workgroups per CU
(F = 2, 4), and the fastest (U, S, F) is written.

Batched launches (csrc/batch_tile_table.inc, rows {n_max, kernel_type,
reach_max, U}): for every kernel, reach class and instance-size bucket, a
batch of instances with n spread +-15 % around the bucket size is timed at
each tile width (pbe_batch_plan_create_ex), and the fastest U is written.
Over a run of sizes, a shape within 3 % of the fastest is kept from the size
below (timing noise then does not flip the table between near-equal shapes).

The reach of a geometric grid of ratio r is log_r(1 / (r - 1)) bins: pairs
further than that below the diagonal are far (both births in (k, k+1)).  The
classes are reach <= 64, 128, 256, 512, 1200 and beyond (tuned at r = 1.06,
1.03, 1.02, 1.01, 1.005, 1.002), plus grids without the geometric lookup
(inv_log_r = 0; tuned on a jittered r = 1.02 grid).  Sizes run while
n ln r <= 330, so that every pivot and every kernel value stays a normal
double; larger n reuses the largest tuned size's row.

Workload: x from 1e-18 and the kernel coefficients of the flocculation
case.  Single launches are timed on the population of the single-instance
throughput workload: lognormal, median x_{n/4}, sigma_ln 0.8, +-5 % noise,
whose tail underflows to zero some 40 sigma out, so that the flushes of
empty bins drop out as they do in practice (the launch forms differ in how
many flushes they make: on a population nonzero everywhere the choice
between near-equal forms can flip).  Batched launches are timed on
N_i = 1e6 r^(-i/2) (1 + 0.05 u_i), nonzero and normal on the whole grid,
so every window flush and atomic is taken.  Each shape is timed over a
hipGraph of back-to-back launches (so small sizes measure the GPU, not the
host's launch rate), best of three replays.

Both tables are written exactly as generated; the library holds no tuning
state at run time.
"""
import argparse
import ctypes
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "python"))

KERNELS = [0, 1, 2, 3, 4, 5]
TILES = [1, 2, 4, 8]
SPLITS = [1, 2, 4, 8, 16]
OPEN = 1 << 30
# (reach_max, ratio the class is timed at); reach_max -1: no geometric lookup
CLASSES = [(64, 1.06), (128, 1.03), (256, 1.02), (512, 1.01), (1200, 1.005), (OPEN, 1.002),
           (-1, 1.02)]
SIZES = [256, 512, 1024, 1536, 2048, 3072, 4096, 6144, 8192, 12288, 16384, 32768, 65536,
         131072]
BATCH_SIZES = [64, 128, 256, 512, 1024, 2048, 4096]
MAX_LOG_SPAN = 330.0
CAPTURE_RELAXED = 2


def grid(n, r, general, seed):
    from pbe_hip import problems as P
    if general:
        return P.jittered_grid(n, 1e-18, r, 0.3, seed)
    return P.geometric_grid(n, 1e-18, r)


def lognormal_populations(x, seed):
    from pbe_hip import problems as P
    n = len(x)
    return P.lognormal_ic(x, x[n // 4], 0.8, 1e12, noise=0.05, seed=seed)


def populations(n, r, seed):
    u = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    return 1e6 * np.exp(-0.5 * np.arange(n) * math.log(r)) * (1.0 + 0.05 * u)


def coefficients(k, x):
    xq = x[len(x) // 4]
    return dict(kernel=k, beta0=1e-17 / (xq ** (2 if k == 2 else 1 if k == 1 else 0)),
                beta_br=6.73e-18, beta_sh=4.0 / 3.0 * 100.0)


class Timer:
    """Mean duration of one call of `launch(stream)` from a captured graph
    of K back-to-back calls (K sized to about 2 ms), best of three replays."""

    def __init__(self, hip):
        self.hip = hip
        self.rt = hip.runtime()
        self.s = hip.Stream()
        self.e0, self.e1 = hip.Event(), hip.Event()

    def __call__(self, launch, target_ms=2.0):
        hip, rt, s = self.hip, self.rt, self.s
        for _ in range(2):
            launch(s)
        s.synchronize()
        self.e0.record(s)
        launch(s)
        self.e1.record(s)
        s.synchronize()
        t1 = max(self.e0.elapsed_ms(self.e1), 1e-3)
        K = int(min(200, max(1, round(target_ms / t1))))
        hip.check(rt.hipStreamBeginCapture(s.handle, CAPTURE_RELAXED), "hipStreamBeginCapture")
        try:
            for _ in range(K):
                launch(s)
        finally:
            g = ctypes.c_void_p()
            hip.check(rt.hipStreamEndCapture(s.handle, ctypes.byref(g)), "hipStreamEndCapture")
        ge = ctypes.c_void_p()
        hip.check(rt.hipGraphInstantiate(ctypes.byref(ge), g, None, None, 0),
                  "hipGraphInstantiate")
        try:
            hip.check(rt.hipGraphLaunch(ge, s.handle), "hipGraphLaunch")
            best = float("inf")
            for _ in range(3):
                self.e0.record(s)
                hip.check(rt.hipGraphLaunch(ge, s.handle), "hipGraphLaunch")
                self.e1.record(s)
                s.synchronize()
                best = min(best, self.e0.elapsed_ms(self.e1) / K)
        finally:
            rt.hipGraphExecDestroy(ge)
            rt.hipGraphDestroy(g)
        return best


def class_sizes(sizes, r, spread=1.0):
    return [n for n in sizes if spread * n * math.log(r) <= MAX_LOG_SPAN]


def tune_single(pbe, hip, timer, sizes, kernels, log):
    """{(kernel, reach_max): [(n, {(U, S, F): t_ms})]}"""
    best = {}
    for ci, (reach_max, r) in enumerate(CLASSES):
        general = reach_max < 0
        for n in class_sizes(sizes, r):
            x = grid(n, r, general, seed=n)
            dN = hip.DeviceArray.from_host(lognormal_populations(x, seed=n + 2))
            dx = hip.DeviceArray.from_host(x)
            drhs = hip.DeviceArray(n)
            drhs.zero()
            line = []
            for k in kernels:
                agg = coefficients(k, x)
                if not general:
                    agg["log_x0"], agg["inv_log_r"] = math.log(x[0]), 1.0 / math.log(x[1] / x[0])
                p = pbe.agg_params(n, agg)

                def run(u, w, f):
                    def launch(s):
                        hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u,
                                                                band_split=w, form=f, stream=s),
                                  f"n {n} kernel {k} tile {u} split {w} form {f}")
                    return timer(launch)
                t = {(u, w, f): run(u, w, f) for u in TILES for w in SPLITS
                     for f in ((1, 2, 3, 4) if u == 4 else (1, 3))}
                bu, bw, bf = min(t, key=t.get)
                best.setdefault((k, reach_max), []).append((n, t))
                line.append(f"k{k}:U{bu}S{bw}F{bf} {t[(bu, bw, bf)] * 1e3:.1f}us")
            log(f"single reach<={reach_max if reach_max >= 0 else 'general'} r={r} n={n}: "
                + "  ".join(line))
    return best


def tune_batch(pbe, hip, timer, sizes, kernels, log):
    """{(kernel, reach_max): [(n, {(U,): t_ms}, pairs)]}"""
    best = {}
    for ci, (reach_max, r) in enumerate(CLASSES):
        general = reach_max < 0
        for nb in class_sizes(sizes, r, spread=1.15):
            count = int(min(4096, max(16, (1 << 27) // (nb * nb))))
            rng = np.random.default_rng(nb)
            ns = np.clip(np.round(nb * rng.uniform(0.85, 1.15, count)).astype(int), 2, None)
            xs, Ns, offs = [], [], []
            off = 0
            for i, n in enumerate(ns):
                xs.append(grid(int(n), r, general, seed=i))
                Ns.append(populations(int(n), r, seed=i + 7))
                offs.append(off)
                off += int(n)
            dN = hip.DeviceArray.from_host(np.concatenate(Ns))
            dx = hip.DeviceArray.from_host(np.concatenate(xs))
            drhs = hip.DeviceArray(off)
            drhs.zero()
            pairs = float(sum(int(n) * (int(n) + 1) // 2 for n in ns))
            line = []
            for k in kernels:
                inst = []
                for n, x, o in zip(ns, xs, offs):
                    c = coefficients(k, x)
                    inst.append(dict(n=int(n), offset=o, beta0=c["beta0"], beta_br=c["beta_br"],
                                     beta_sh=c["beta_sh"],
                                     log_x0=0.0 if general else math.log(x[0]),
                                     inv_log_r=0.0 if general else 1.0 / math.log(x[1] / x[0])))
                t = {}
                for u in TILES:
                    plan = pbe.BatchPlan(inst, k, -1, 0, tile_u=u)

                    def launch(s, plan=plan):
                        hip.check(plan.launch(dN, dx, None, None, drhs, stream=s),
                                  f"batch {nb} kernel {k} tile {u}")
                    t[(u,)] = timer(launch)
                    plan.destroy()
                bu = min(t, key=t.get)
                best.setdefault((k, reach_max), []).append((nb, t, pairs))
                line.append(f"k{k}:U{bu[0]} {pairs / (t[bu] * 1e-3):.2e}/s")
            log(f"batch reach<={reach_max if reach_max >= 0 else 'general'} r={r} n~{nb} "
                f"x{count}: " + "  ".join(line))
    return best


def choices(ent, tol=0.03):
    """Shape per size, smallest first: the fastest, unless the previous
    size's shape is within tol of it (near-ties then keep one shape over a
    run of sizes instead of flipping on timing noise)."""
    out, prev = [], None
    for e in sorted(ent, key=lambda e: e[0]):
        t = e[1]
        fast = min(t, key=t.get)
        pick = prev if prev is not None and t[prev] <= (1.0 + tol) * t[fast] else fast
        out.append((e[0], pick, t[pick]))
        prev = pick
    return out


def merged_rows(best):
    """Rows of one table: per (kernel, class), consecutive sizes with the same
    shape merged; each size covers up to the midpoint to the next one and the
    largest up to 2^30."""
    rows = []
    for k in KERNELS:
        for reach_max, _ in CLASSES:
            ent = choices(best.get((k, reach_max), []))
            for i, (n, shape, _) in enumerate(ent):
                nmax = (n + ent[i + 1][0]) // 2 if i + 1 < len(ent) else OPEN
                if rows and rows[-1][1:3] == (k, reach_max) and rows[-1][3:] == shape:
                    rows[-1] = (nmax, k, reach_max) + shape
                else:
                    rows.append((nmax, k, reach_max) + shape)
    return rows


def class_name(reach_max):
    return "no lookup" if reach_max < 0 else ("any" if reach_max == OPEN else f"<= {reach_max}")


def summary_lines(best, batch):
    """One markdown row per (reach class, n): each kernel's chosen shape and
    its time (single launches) or pair rate (batched launches)."""
    ks = sorted({k for k, _ in best})
    cells = {}
    for (k, c), ent in best.items():
        for i, (n, shape, t) in enumerate(choices(ent)):
            pairs = ent[[e[0] for e in ent].index(n)][2] if batch else None
            val = (f"U{shape[0]} {pairs / (t * 1e-3):.2e}/s" if batch
                   else f"U{shape[0]} S{shape[1]} F{shape[2]} {t * 1e3:.1f}")
            cells[(c, n, k)] = val
    head = "| reach (bins) | n" + (" ~" if batch else "") + " | " + " | ".join(
        f"kernel {k}" for k in ks) + " |"
    lines = [head, "|" + "---|" * (2 + len(ks))]
    for reach_max, r in CLASSES:
        for n in sorted({n for (c, n, _) in cells if c == reach_max}):
            lines.append(f"| {class_name(reach_max)} (r = {r}) | {n} | " + " | ".join(
                cells.get((reach_max, n, k), "") for k in ks) + " |")
    return lines


def write_table(path, header, rows):
    with open(path, "w") as f:
        for h in header:
            f.write("// " + h + "\n")
        for r in rows:
            f.write("{" + ", ".join(str(v) for v in r) + "},\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--part", choices=["single", "batch", "both"], default="both")
    ap.add_argument("--out-dir", default=os.path.join(HERE, "csrc"))
    ap.add_argument("--summary", default=None, help="markdown summary of the timings")
    ap.add_argument("--quick", action="store_true", help="few sizes and kernels (a smoke run)")
    args = ap.parse_args()
    import pbe_hip as pbe
    from pbe_hip import hip
    timer = Timer(hip)
    t0 = time.time()

    def log(s):
        print(f"[{time.time() - t0:6.0f}s] {s}", flush=True)

    sizes = [256, 4096] if args.quick else SIZES
    bsizes = [128, 1024] if args.quick else BATCH_SIZES
    kernels = [0, 5] if args.quick else KERNELS
    summary = []
    if args.part in ("single", "both"):
        best = tune_single(pbe, hip, timer, sizes, kernels, log)
        write_table(os.path.join(args.out_dir, "tile_table.inc"),
                    ["{n_max, kernel_type, reach_max (-1: grids without the geometric",
                     "lookup), columns per lane, band-tile row slices, launch form}; first",
                     "match wins.  Generated by pbe-fixed-pivot-cuda_amd/autotune.py on an",
                     "MI355X (gfx950); do not edit."],
                    merged_rows(best))
        summary += ["Single launches: chosen U (columns per lane), S (band split), F "
                    "(launch form: 1 two launches, 2 fused band + strips, 3 one tile launch, "
                    "4 fused at three workgroups per CU) "
                    "and the time of one aggregation call in us.", ""]
        summary += summary_lines(best, batch=False)
    if args.part in ("batch", "both"):
        bbest = tune_batch(pbe, hip, timer, bsizes, kernels, log)
        write_table(os.path.join(args.out_dir, "batch_tile_table.inc"),
                    ["{n_max, kernel_type, reach_max (-1: grids without the geometric",
                     "lookup), columns per lane}; first match wins.  Generated by",
                     "pbe-fixed-pivot-cuda_amd/autotune.py on an MI355X (gfx950); do not",
                     "edit."],
                    merged_rows(bbest))
        summary += ["", "Batched launches: chosen U and the pair rate (pair updates per "
                    "second) of one batched aggregation call.", ""]
        summary += summary_lines(bbest, batch=True)
    if args.summary:
        with open(args.summary, "w") as f:
            f.write("\n".join(summary) + "\n")
    log("done")


if __name__ == "__main__":
    main()
