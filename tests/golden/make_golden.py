"""Generate tests/golden/fixed_pivot_golden.json with mpmath (50 digits).

An independent restatement of the fixed-pivot source terms, written from the
mathematical definition rather than from the pair-loop code: with the
Kumar-Ramkrishna hat functions eta_i(v) (docs/theory.md sections 3-5 of the
reference), clipped to the end pivots outside [x_0, x_{n-1}],

  aggregation  R_i = sum_{j<=k} (1 - delta_jk/2) beta_jk N_j N_k eta_i(v_jk)
                     - N_i sum_k beta_ik N_k
  breakage     R_i = - [S_i N_i]_g + sum_j [S_j N_j]_g sum_q bw_q eta_i(v_jq)

where [.]_g applies the reference's guard (S != 0 and N > 2.2e-308,
src/breakage_kernels.cuh:131,195).  The event volumes v_jk = x_j + x_k and
v_jq = x_j t_q are the double-precision values the reference forms
(src/aggregation_kernels.cuh:190, src/breakage_kernels.cuh:199); everything
after that is exact to 50 digits, including the cube roots and powers.
Besides rhs the fixture stores A_i = sum of |contributions| per bin, the
scale of the comparison.

    python tests/golden/make_golden.py      # rewrites the JSON next to it
"""
import json
import math
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "fixed_pivot_golden.json")


def beta_exact(kernel, xj, xk, b0, bbr, bsh):
    xj, xk = mp.mpf(xj), mp.mpf(xk)
    cj, ck = mp.cbrt(xj), mp.cbrt(xk)
    br = mp.mpf(bbr) * (cj / ck + ck / cj + 2)
    sh = mp.mpf(bsh) * (cj + ck) ** 3
    return {0: mp.mpf(b0), 1: mp.mpf(b0) * (xj + xk), 2: mp.mpf(b0) * xj * xk,
            3: br, 4: sh, 5: br + sh}[kernel]


def hat_weights(x, v):
    """[(i, eta_i(v))] for the fixed-pivot hat functions, clipped at the ends."""
    n = len(x)
    if v >= x[-1]:
        return [(n - 1, mp.mpf(1))]
    if v <= x[0]:
        return [(0, mp.mpf(1))]
    # the pivot interval [x_i, x_{i+1}) containing v, by a linear scan
    i = max(m for m in range(n - 1) if x[m] <= v)
    xl, xh, vv = mp.mpf(x[i]), mp.mpf(x[i + 1]), mp.mpf(v)
    up = (vv - xl) / (xh - xl)
    return [(i, 1 - up), (i + 1, up)]


def aggregation_golden(x, N, kernel, b0, bbr, bsh):
    n = len(x)
    R = [mp.mpf(0)] * n
    A = [mp.mpf(0)] * n
    for i in range(n):                       # death: full row sum
        d = mp.mpf(N[i]) * mp.fsum(beta_exact(kernel, x[i], x[k], b0, bbr, bsh) * mp.mpf(N[k])
                                   for k in range(n))
        R[i] -= d
        for k in range(n):                   # |contributions| as the pair loop forms them
            c = beta_exact(kernel, x[i], x[k], b0, bbr, bsh) * mp.mpf(N[i]) * mp.mpf(N[k])
            A[i] += abs(c)
    for j in range(n):                       # birth: upper triangle
        for k in range(j, n):
            rate = beta_exact(kernel, x[j], x[k], b0, bbr, bsh) * mp.mpf(N[j]) * mp.mpf(N[k])
            if j == k:
                rate /= 2
            v = float(x[j] + x[k])
            for i, eta in hat_weights(x, v):
                R[i] += eta * rate
                A[i] += abs(eta * rate)
    return R, A


def selection_exact(sel, v, S0, v_ref, alpha, v_min):
    v = mp.mpf(v)
    if sel == 0:
        return mp.mpf(S0)
    if sel == 1:
        return mp.mpf(S0) * v / mp.mpf(v_ref)
    if sel == 2:
        return mp.mpf(S0) * mp.power(v / mp.mpf(v_ref), mp.mpf(alpha))
    return mp.mpf(S0) if v > v_min else mp.mpf(0)


def breakage_golden(x, N, t_q, bw_q, sel, S0, v_ref, alpha, v_min):
    n = len(x)
    R = [mp.mpf(0)] * n
    A = [mp.mpf(0)] * n
    for j in range(n):
        S = selection_exact(sel, x[j], S0, v_ref, alpha, v_min)
        if not (S != 0 and N[j] > 2.2e-308):
            continue
        sn = S * mp.mpf(N[j])
        R[j] -= sn
        A[j] += abs(sn)
        for t, bw in zip(t_q, bw_q):
            v = float(x[j] * t)
            c = sn * mp.mpf(bw)
            for i, eta in hat_weights(x, v):
                R[i] += eta * c
                A[i] += abs(eta * c)
    return R, A


def geometric(n, x0, r):
    x = [x0]
    for _ in range(1, n):
        x.append(x[-1] * r)
    return x


def jittered(n, b0, r, jit, seed):
    rng = np.random.default_rng(seed)
    d = rng.uniform(-jit, jit, n + 1)
    b = b0 * np.power(r, np.arange(n + 1) + d)
    return [float(v) for v in np.sqrt(b[:-1] * b[1:])]


def gl(m, a, b):
    t, w = np.polynomial.legendre.leggauss(m)
    return list(0.5 * (b - a) * t + 0.5 * (b + a)), list(0.5 * (b - a) * w)


def daughters():
    tu, wu = gl(4, 0.0, 1.0)
    t1, w1 = gl(2, 0.0, 0.5)
    t2, w2 = gl(2, 0.5, 1.0)
    tp, wp = t1 + t2, w1 + w2
    return {
        "symmetric": ([0.5], [2.0]),
        "erosion": ([0.95, 0.05], [1.0, 1.0]),
        "uniform4": (tu, [2.0 * w for w in wu]),
        "powerlaw2_4": (tp, [4.0 * t * t * w for t, w in zip(tp, wp)]),
    }


def main():
    rng = np.random.default_rng(20261203)
    grids = {
        "geometric8": (geometric(8, 1e-18, 1.7), True),
        "jittered7": (jittered(7, 1e-18, 1.6, 0.2, 5), False),
        "coarse5": (geometric(5, 1e-18, 3.0), True),
        "single1": ([2e-18], False),
        "pair2": (geometric(2, 1e-18, 1.5), True),
    }
    cases = []
    for gname, (x, geo) in grids.items():
        n = len(x)
        N = list(rng.uniform(1e10, 1e12, n))
        if n > 3:
            N[1] = 0.0
        if gname == "jittered7":
            N[3] = -2.5e11          # aggregation has no population guard
        log_x0 = math.log(x[0])
        inv_log_r = 1.0 / math.log(x[1] / x[0]) if (geo and n > 1) else 0.0
        for kernel in range(6):
            b0 = 1e-17 if kernel != 2 else 1e-17 / x[n // 2]
            if kernel == 1:
                b0 = 1e-17 / x[n // 2]
            bbr, bsh = 6.73e-18, 2.0e-6
            R, A = aggregation_golden(x, N, kernel, b0, bbr, bsh)
            cases.append(dict(kind="aggregation", grid=gname, x=x, N=N, kernel=kernel,
                              beta0=b0, beta_br=bbr, beta_sh=bsh, log_x0=log_x0,
                              inv_log_r=inv_log_r,
                              rhs=[float(v) for v in R], A=[float(v) for v in A]))
    for gname in ("geometric8", "jittered7", "coarse5"):
        x = grids[gname][0]
        n = len(x)
        N = list(rng.uniform(1e10, 1e12, n))
        N[2] = 0.0
        N[n - 1] = 1e-310           # below the breakage guard
        for dname, (t_q, bw_q) in daughters().items():
            for sel in range(4):
                v_ref, alpha, v_min = x[n // 2], 1.0 / 3.0, x[n // 2]
                R, A = breakage_golden(x, N, t_q, bw_q, sel, 1e-3, v_ref, alpha, v_min)
                cases.append(dict(kind="breakage", grid=gname, daughter=dname, x=x, N=N,
                                  t_q=[float(t) for t in t_q], bw_q=[float(b) for b in bw_q],
                                  selection=sel, S0=1e-3, v_ref=v_ref, alpha=alpha,
                                  v_min=v_min, rhs=[float(v) for v in R],
                                  A=[float(v) for v in A]))
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py (mpmath, 50 digits)",
                   "cases": cases}, f, indent=0)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
