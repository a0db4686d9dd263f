"""Closed forms of the discrete fixed-pivot system used as trajectory pins.

Each RHS conserves M1 = sum x_i N_i exactly while nothing clips at the grid
ends, so every RK4 stage does, and the zeroth moment M0 = sum N_i follows a
scalar ODE that RK4 then integrates with its own polynomial:
  sum kernel  b0 (x_j + x_k):   dM0/dt = -b0 M0 M1   -> M0 R(-b0 M1 dt)^K,
                                R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24
  product     b0 x_j x_k:       dM0/dt = -b0 M1^2 / 2 (linear in t: exact)
  linear selection S0 v/v_ref, symmetric binary daughters (two halves):
                                dM0/dt = S0 M1 / v_ref (linear in t: exact)
(Golovin, product-kernel and Ziff-McGrady moment laws; the identities hold
for the discretisation itself, not just its continuum limit.)
"""
import numpy as np

from pbe_hip import problems as P

N_BINS, DT, STEPS = 256, 1.0, 1000


def grid():
    x = P.geometric_grid(N_BINS, 1e-18, 2.0 ** 0.25)
    return x, P.grid_lookup(x)


def sum_case():
    x, (l0, ilr) = grid()
    N0 = P.exponential_ic_cells(x, 1e12, 10.0 * x[0])
    M0, M1 = N0.sum(), (x * N0).sum()
    b0 = 1.0 / (M1 * 1000.0)
    agg = dict(kernel=1, beta0=b0, beta_br=0.0, beta_sh=0.0, log_x0=l0, inv_log_r=ilr)
    z = -b0 * M1 * DT
    R = 1.0 + z + z * z / 2.0 + z ** 3 / 6.0 + z ** 4 / 24.0
    return x, N0, dict(agg=agg), M0 * R ** STEPS, M1


def product_case():
    x, (l0, ilr) = grid()
    N0 = P.exponential_ic_cells(x, 1e12, 10.0 * x[0])
    M0, M1, M2 = N0.sum(), (x * N0).sum(), (x * x * N0).sum()
    b0 = 0.3 / (M2 * 1000.0)   # 0.3 of the way to gelation at t = 1000
    agg = dict(kernel=2, beta0=b0, beta_br=0.0, beta_sh=0.0, log_x0=l0, inv_log_r=ilr)
    return x, N0, dict(agg=agg), M0 - 0.5 * b0 * M1 * M1 * DT * STEPS, M1


def breakage_case():
    x, _ = grid()
    N0 = P.lognormal_ic(x, x[160], 0.5, 1e12)
    M0, M1 = N0.sum(), (x * N0).sum()
    t_q, bw_q = P.daughter_symmetric()
    brk = dict(selection=1, S0=1e-3, v_ref=x[160], alpha=1.0, v_min=0.0)
    return (x, N0, dict(brk=brk, t_q=t_q, bw_q=bw_q),
            M0 + 1e-3 * M1 / x[160] * DT * STEPS, M1)


CASES = {"sum": sum_case, "product": product_case, "breakage": breakage_case}


def check(x, N, M0_expected, M1, tol_m0, tol_m1=1e-12):
    m0 = N.sum()
    assert np.all(np.isfinite(N))
    assert abs(m0 - M0_expected) <= tol_m0 * abs(M0_expected), (m0, M0_expected)
    assert abs((x * N).sum() - M1) <= tol_m1 * M1
