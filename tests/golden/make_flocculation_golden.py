"""Generate tests/golden/flocculation_two_stage.json: the oracle's RK4 run of
the reference's two-stage flocculation case study at N = 960
(examples/flocculation_two_stage/main.cpp: 3600 steps of 0.5 s, the shear
rate switched from 100 to 20 1/s after the step that starts before
t = 600 s), with d43 at t_mid and t_end, the final M0, the first-moment drift
and the final state.  The GPU example program
(pbe-fixed-pivot-cuda_amd/examples/flocculation_two_stage) is checked
against it.

    python tests/golden/make_flocculation_golden.py    # about a minute on 8 cores
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "python"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
OUT = os.path.join(HERE, "flocculation_two_stage.json")


def run(n=960, threads=8):
    import oracle
    from pbe_hip import problems as P
    c = P.config_flocculation_two_stage(n)
    x, N = c["x"], c["N0"].copy()
    kw_hi = dict(agg=c["agg_high"], brk=c["brk_high"], t_q=c["t_q"], bw_q=c["bw_q"],
                 threads=threads)
    kw_lo = dict(agg=c["agg_low"], brk=c["brk_low"], t_q=c["t_q"], bw_q=c["bw_q"],
                 threads=threads)
    M1_0 = float(np.dot(x, N))
    t = time.time()
    # to t_mid (the state at step mid_step), the rest of the high-shear steps,
    # then the low-shear steps
    N = oracle.rk4(N, x, c["dt"], c["mid_step"], **kw_hi)
    d43_mid = P.d43(N, x, c["d0"]) * 1e6
    N = oracle.rk4(N, x, c["dt"], c["n_high"] - c["mid_step"], **kw_hi)
    N = oracle.rk4(N, x, c["dt"], c["steps"] - c["n_high"], **kw_lo)
    print(f"oracle RK4, {c['steps']} steps: {time.time() - t:.0f} s", flush=True)
    return dict(n=n, steps=c["steps"], steps_high=c["n_high"], mid_step=c["mid_step"],
                d43_mid_um=d43_mid, d43_end_um=P.d43(N, x, c["d0"]) * 1e6,
                M0_end=float(sum(N)), M1_drift=abs(float(np.dot(x, N)) - M1_0) / M1_0,
                N_end=N.tolist())


if __name__ == "__main__":
    doc = {"about": "oracle RK4 of the two-stage flocculation case study "
                    "(tests/golden/make_flocculation_golden.py)"}
    doc.update(run())
    print({k: v for k, v in doc.items() if k != "N_end"})
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
    print("wrote", OUT)
