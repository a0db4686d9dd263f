"""Pick the aggregation strip shape per (n, kernel) by measurement.

    python pbe-fixed-pivot-cuda_amd/autotune.py --out tile_table.inc [--json results.json]

For each size bucket and collision kernel, every block width (64*U columns,
U = 1, 2, 4, 8 columns per lane) and band split S = 1, 2, 4, 8, 16 (row
slices per diagonal-band tile; offset segments per block for the band
sweeps of 256-column blocks) is timed with HIP events over
back-to-back launches on the flocculation workload (lognormal N, geometric
grid r = 1.02), and the fastest (U, S) is written as a rule {n_max, kernel,
U, S} in the C include that the library compiles in (csrc/tile_table.inc).  The table is read-only
data built offline: the library holds no tuning state at run time.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "python"))

SIZES = [256, 512, 1024, 2048, 4096, 8192, 16384]
KERNELS = [0, 1, 2, 3, 4, 5]
TILES = [1, 2, 4, 8]
WEIGHTS = [1, 2, 4, 8, 16]


def time_launch(pbe, hip, N, x, agg, u, w, reps):
    n = len(N)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    drhs.zero()
    p = pbe.agg_params(n, agg)
    s = hip.Stream()

    def launch():
        hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u, band_split=w,
                                                stream=s), f"tile {u} split {w}")
    for _ in range(3):
        launch()
    s.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    for _ in range(reps):
        launch()
    e1.record(s)
    s.synchronize()
    return e0.elapsed_ms(e1) / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "tile_table.inc"))
    ap.add_argument("--json", default=None)
    ap.add_argument("--sizes", default=None)
    args = ap.parse_args()
    import pbe_hip as pbe
    from pbe_hip import hip, problems as P
    sizes = [int(s) for s in args.sizes.split(",")] if args.sizes else SIZES
    results = {}
    best = {}
    for n in sizes:
        x, N, agg0, _, _, _ = P.config_flocculation(n)
        for k in KERNELS:
            agg = dict(agg0, kernel=k, beta0=1e-17 / (x[n // 4] ** (2 if k == 2 else 1 if k == 1 else 0)))
            reps = max(5, min(200, int(4e9 / (n * n))))
            t = {(u, w): time_launch(pbe, hip, N, x, agg, u, w, reps)
                 for u in TILES for w in WEIGHTS}
            results[f"{n}/{k}"] = {f"U{u}W{w}": v for (u, w), v in t.items()}
            best[(n, k)] = min(t, key=t.get)
            bu, bw = best[(n, k)]
            print(f"n={n:6d} kernel={k} " + " ".join(
                f"U{u}:{min(t[(u, w)] for w in WEIGHTS) * 1e3:8.1f}us" for u in TILES)
                + f"  -> U={bu} W={bw} {t[(bu, bw)] * 1e3:.1f}us", flush=True)
    lines = ["// {n_max, kernel_type (-1 = any), columns per lane, band-tile row slices};",
             "// first match wins.  Generated by pbe-fixed-pivot-cuda_amd/autotune.py on",
             "// an MI355X (gfx950)."]
    for k in KERNELS:
        prev = None
        for i, n in enumerate(sizes):
            u, w = best[(n, k)]
            nmax = (sizes[i] + sizes[i + 1]) // 2 if i + 1 < len(sizes) else (1 << 30)
            if prev is not None and prev[1] == (u, w):
                lines[-1] = f"{{{nmax}, {k}, {u}, {w}}},"
            else:
                lines.append(f"{{{nmax}, {k}, {u}, {w}}},")
            prev = (nmax, (u, w))
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(results, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
