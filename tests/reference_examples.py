"""The reference's five example programs as known-answer cases.

Each case restates an example's grid, initial condition, kernel, integrator
and error measure, and carries the error the reference documents for it
(docs/examples.md).  Grids and initial conditions follow
examples/common/pbe_examples_utils.hpp: make_geometric_grid (:83-90, repeated
multiplication), make_geometric_grid_range (:93-99), make_exponential_ic
(:121-136, exact cell integrals); the analytic moments are those of
:198-252 (constant, sum and product kernels) and :263-280 (Ziff-McGrady).

Documented figures reproduced to their two printed digits (within 5 %):
example 1 N_tot 1.7e-4; example 2 M0 1.4e-4, M1 1.6e-4; example 3 M0 1.7e-4,
M2 4.8e-3, M1 1.6e-4; example 5 M0 2.2e-4.  Two documented figures are not
what the example computes, and the cases pin the exact behaviour instead:
example 4's N_tot error is documented as ~1e-4, but with linear selection
M0 grows exactly linearly under explicit Euler, so the error is at rounding
level; example 5's M1 error is documented as 1.8e-4, but the initial
condition on that grid already differs from N0 vc by 2.44e-4 and the
breakage conserves M1 (drift ~1e-9, from fragments clipped into bin 0).
"""
import math

import numpy as np

# Gauss-Legendre nodes and weights on (0, 1) as listed in
# examples/breakage_linear_uniform/main.cpp (8 points)
GL8_T = np.array([0.019855071751232, 0.101666761293187, 0.237233795041836,
                  0.408282678752175, 0.591717321247825, 0.762766204958164,
                  0.898333238706813, 0.980144928248768])
GL8_W = np.array([0.050614268145188, 0.111190517226687, 0.156853322938943,
                  0.181341891689181, 0.181341891689181, 0.156853322938943,
                  0.111190517226687, 0.050614268145188])


def geometric_grid(n, v_min, r):
    x = np.empty(n)
    x[0] = v_min
    for i in range(1, n):
        x[i] = x[i - 1] * r
    return x


def geometric_grid_range(n, v_min, v_max):
    return geometric_grid(n, v_min, (v_max / v_min) ** (1.0 / (n - 1)))


def exponential_ic(x, N0, vc):
    sr = math.sqrt(x[1] / x[0])
    return N0 * (np.exp(-(x / sr) / vc) - np.exp(-(x * sr) / vc))


def lookup(x):
    return math.log(x[0]), 1.0 / math.log(x[1] / x[0])


def rel(a, b):
    return abs(a - b) / abs(b)


def example(k):
    """Case k (1-5): dict with x, N0, method ('euler' | 'rk4'), dt, steps,
    the RHS keywords (agg / brk, t_q, bw_q) and check(N) -> {name: error}
    with `documented` {name: figure} and `exact` {name: bound}."""
    if k == 1:   # examples/aggregation_constant: monodisperse IC, Euler
        x = geometric_grid(256, 1e-18, 1.122018)
        N0 = np.zeros(256)
        N0[0] = 1e14
        l0, ilr = lookup(x)
        M1 = float((x * N0).sum())
        exact_N = 1e14 / (1.0 + 2000.0 / (2.0 / (1e-17 * 1e14)))
        return dict(x=x, N0=N0, method="euler", dt=1.0, steps=2000,
                    kw=dict(agg=dict(kernel=0, beta0=1e-17, beta_br=0.0, beta_sh=0.0,
                                     log_x0=l0, inv_log_r=ilr)),
                    check=lambda N: {"N_tot": rel(N.sum(), exact_N),
                                     "M1": rel((x * N).sum(), M1)},
                    documented={"N_tot": 1.7e-4}, exact={"M1": 1e-12})
    if k in (2, 3):   # examples/aggregation_sum, aggregation_product: RK4
        x = geometric_grid_range(300, 1e-4, 1e4)
        N0 = exponential_ic(x, 1.0, 1.0)
        l0, ilr = lookup(x)
        t = 0.3
        agg = dict(kernel=1 if k == 2 else 2, beta0=1.0, beta_br=0.0, beta_sh=0.0,
                   log_x0=l0, inv_log_r=ilr)
        if k == 2:
            check = lambda N: {"M0": rel(N.sum(), math.exp(-t)),
                               "M1": rel((x * N).sum(), 1.0)}
            documented = {"M0": 1.4e-4, "M1": 1.6e-4}
        else:
            M2 = 2.0 / (1.0 - 2.0 * t)
            check = lambda N: {"M0": rel(N.sum(), 1.0 - 0.5 * t),
                               "M2": rel((x * x * N).sum(), M2),
                               "M1": rel((x * N).sum(), 1.0)}
            documented = {"M0": 1.7e-4, "M2": 4.8e-3, "M1": 1.6e-4}
        return dict(x=x, N0=N0, method="rk4", dt=1e-4, steps=3000, kw=dict(agg=agg),
                    check=check, documented=documented, exact={})
    if k == 4:   # examples/breakage_linear_symmetricbinary: Euler
        x = geometric_grid(512, 1e-18, 1.122018)
        N0 = np.zeros(512)
        N0[384] = 1e14
        V = float((x * N0).sum())
        brk = dict(selection=1, S0=1e-3, v_ref=x[384], alpha=1.0, v_min=0.0)
        return dict(x=x, N0=N0, method="euler", dt=0.01, steps=10000,
                    kw=dict(brk=brk, t_q=np.array([0.5]), bw_q=np.array([2.0])),
                    check=lambda N: {"N_tot": rel(N.sum(), 1e14 * (1.0 + 1e-3 * 100.0)),
                                     "V_tot": rel((x * N).sum(), V)},
                    documented={}, exact={"N_tot": 1e-12, "V_tot": 1e-12})
    if k == 5:   # examples/breakage_linear_uniform (Ziff-McGrady): Euler
        x = geometric_grid_range(200, 1e-4, 1e1)
        N0 = exponential_ic(x, 1.0, 1.0)
        M1_0 = float((x * N0).sum())
        brk = dict(selection=1, S0=1.0, v_ref=1.0, alpha=1.0, v_min=0.0)
        return dict(x=x, N0=N0, method="euler", dt=1e-3, steps=5000,
                    kw=dict(brk=brk, t_q=GL8_T, bw_q=2.0 * GL8_W),
                    check=lambda N: {"M0": rel(N.sum(), 6.0),
                                     "M1_drift": rel((x * N).sum(), M1_0),
                                     "M1": rel((x * N).sum(), 1.0)},
                    documented={"M0": 2.2e-4}, exact={"M1_drift": 1e-8, "M1": 1e-3})
    raise ValueError(k)


def assert_matches(case, N):
    err = case["check"](N)
    for name, fig in case["documented"].items():
        assert abs(err[name] - fig) <= 0.05 * fig, (name, err[name], fig)
    for name, bound in case["exact"].items():
        assert err[name] <= bound, (name, err[name], bound)
    return err
