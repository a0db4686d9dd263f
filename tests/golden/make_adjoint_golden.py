"""Generate tests/golden/adjoint_config5_cs.json: complex-step gradients of
the config-5 parameter-fit loss, for the GPU adjoint test at full size.

Workload: n = 1024 geometric grid r = 1.03 from x_0 = 1e-18, Brownian +
shear aggregation and power-law breakage (exponent 1/3) with symmetric-binary
daughters (pbe_hip.problems.config_adjoint), explicit RK4 with dt = 0.5 s for
2000 steps.  Loss: L = sum_i m_i (log max(N_i(T), eps) - log T_i)^2 against a
target PSD T from the same solve with the four fitted parameters (beta_br,
beta_sh, S0, alpha) each multiplied by U[0.5, 2] (numpy default_rng(2)).

The derivatives come from the oracle's complex-step RK4 (oracle/pbe_oracle_cs.c,
itself pinned by tests/test_oracle_cs.py): dL/dtheta_p = Im L(theta + i h e_p)/h
with h = 1e-30, and the same along one direction dN0 of the initial state.
These are the derivatives of the discrete RK4 map to machine precision.

    python tests/golden/make_adjoint_golden.py    # about 7 min on 8 cores
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
OUT = os.path.join(HERE, "adjoint_config5_cs.json")

N_BINS, DT, STEPS, SEED_TARGET, SEED_DIR = 1024, 0.5, 2000, 2, 3


def main(threads=8):
    import oracle
    from pbe_hip import problems as P
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(N_BINS)
    th0 = np.array([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]])
    th_true = th0 * np.random.default_rng(SEED_TARGET).uniform(0.5, 2.0, 4)
    a_t, b_t = P.with_params(agg, brk, th_true)
    t = time.time()
    T = oracle.cs_rk4_final(N0, x, DT, STEPS, a_t, b_t, t_q, bw_q, threads=threads)
    print(f"target solve {time.time() - t:.0f} s", flush=True)
    eps = 1e-30 * float(N0.sum())
    logT = np.log(np.maximum(T, eps))
    mask = (T > 1e-8 * T.max()).astype(float)
    grads = {}
    L = None
    for p, slot in enumerate((1, 2, 3, 4)):
        e = np.zeros(5)
        e[slot] = 1.0
        t = time.time()
        Lp, dL = oracle.cs_loss(N0, x, DT, STEPS, logT, mask, eps, agg, brk, t_q, bw_q,
                                dtheta=e, threads=threads)
        print(f"{P.FIT_PARAMS[p]}: L = {Lp:.17g} dL = {dL:.17g} ({time.time() - t:.0f} s)",
              flush=True)
        L = Lp if L is None else L
        assert abs(Lp - L) <= 1e-13 * abs(L)   # threads sum in varying order
        grads[P.FIT_PARAMS[p]] = dL
    dN0 = np.random.default_rng(SEED_DIR).normal(size=N_BINS) * N0
    _, dL_dir = oracle.cs_loss(N0, x, DT, STEPS, logT, mask, eps, agg, brk, t_q, bw_q,
                               dN0=dN0, threads=threads)
    print(f"dN0 direction: {dL_dir:.17g}", flush=True)
    doc = {
        "about": "complex-step gradient of the config-5 RK4 log-misfit loss "
                 "(tests/golden/make_adjoint_golden.py)",
        "n": N_BINS, "dt": DT, "steps": STEPS,
        "theta0": th0.tolist(), "theta_true": th_true.tolist(),
        "eps": eps, "logT": logT.tolist(), "mask": mask.tolist(),
        "L": L, "grad": grads,
        "dN0": dN0.tolist(), "grad_dN0": dL_dir,
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
