"""CPU tests that pin the complex-step oracle (oracle/pbe_oracle_cs.c) before
it is trusted as the checker of the hand-written VJP and RK4 adjoint.

Each pin is chosen so that a slip in the complex restatement fails it:
  * its real part is the right-hand side of the real oracle (which the golden
    vectors and identities of test_oracle.py pin), entry by entry;
  * R is linear in beta0, beta_br, beta_sh and S0, so dR/dcoef is R evaluated
    with that coefficient 1 and the others 0 -- a closed form the real oracle
    gives directly (a wrong or dropped imaginary part fails it);
  * R is quadratic in N (aggregation) plus linear (breakage), so a central
    difference (R(N + d) - R(N - d)) / 2 is the exact directional derivative,
    up to rounding only;
  * d/dalpha of the power-law selection and the RK4 misfit loss against
    central differences of the real oracle's RK4 (h = 1e-5 relative).
"""
import numpy as np
import pytest

from parity import assert_parity


def _case(n=160, kernel=5, selection=2, daughter="uniform", nq=8):
    from pbe_hip import problems as P
    x, N, agg, brk, _, _ = P.config_adjoint(n)
    agg = dict(agg, kernel=kernel, beta0=1e-17)
    if kernel == 1:
        agg["beta0"] = 1e-17 / x[n // 2]
    if kernel == 2:
        agg["beta0"] = 1e-17 / x[n // 2] ** 2
    brk = dict(brk, selection=selection, v_min=x[n // 2], alpha=0.4)
    t_q, bw_q = P.daughter_table(daughter, nq, q=2.0, eps=0.05)
    return x, N, agg, brk, t_q, bw_q


def _real_rhs(orc, N, x, agg, brk, t_q, bw_q):
    r = np.zeros(len(N))
    A = np.zeros(len(N))
    if agg:
        ra, Aa, _ = orc.aggregation_rhs(N, x, **agg)
        r, A = r + ra, A + Aa
    if brk:
        rb, Ab, _ = orc.breakage_rhs(N, x, t_q, bw_q, **brk)
        r, A = r + rb, A + Ab
    return r, A


@pytest.mark.parametrize("kernel,selection,daughter", [
    (5, 2, "uniform"), (0, 0, "symmetric"), (1, 1, "erosion"), (2, 3, "powerlaw"),
    (3, 2, "symmetric"), (4, 1, "uniform")])
def test_cs_real_part_is_the_oracle(oracle_mod, kernel, selection, daughter):
    x, N, agg, brk, t_q, bw_q = _case(kernel=kernel, selection=selection, daughter=daughter)
    ref, A = _real_rhs(oracle_mod, N, x, agg, brk, t_q, bw_q)
    re, _ = oracle_mod.cs_rhs(N, x, agg, brk, t_q, bw_q, threads=2)
    assert_parity(re, ref, A, tol=1e-13, what="cs real part")


@pytest.mark.parametrize("kernel,slot", [(0, 0), (1, 0), (2, 0), (3, 1), (4, 2), (5, 1), (5, 2)])
def test_cs_coefficient_derivative_is_linear_closed_form(oracle_mod, kernel, slot):
    """dR/dbeta for the coefficient in `slot` equals R_agg with that
    coefficient 1 and the others 0."""
    x, N, agg, brk, t_q, bw_q = _case(kernel=kernel)
    e = np.zeros(5)
    e[slot] = 1.0
    _, dR = oracle_mod.cs_rhs(N, x, agg, None, dtheta=e, threads=2)
    unit = dict(agg, beta0=0.0, beta_br=0.0, beta_sh=0.0)
    unit[("beta0", "beta_br", "beta_sh")[slot]] = 1.0
    ref, A, _ = oracle_mod.aggregation_rhs(N, x, **unit)
    assert np.any(ref != 0.0)
    assert_parity(dR, ref, A, tol=1e-13, what=f"d/dcoef{slot}")
    # a coefficient the kernel does not use has a zero derivative
    users = {0: (0, 1, 2), 1: (3, 5), 2: (4, 5)}
    for other in range(3):
        if kernel not in users[other]:
            f = np.zeros(5)
            f[other] = 1.0
            _, d0 = oracle_mod.cs_rhs(N, x, agg, None, dtheta=f, threads=2)
            assert np.all(d0 == 0.0), (kernel, other)


@pytest.mark.parametrize("selection", range(4))
def test_cs_s0_derivative_is_linear_closed_form(oracle_mod, selection):
    x, N, agg, brk, t_q, bw_q = _case(selection=selection, daughter="powerlaw")
    e = np.zeros(5)
    e[3] = 1.0
    _, dR = oracle_mod.cs_rhs(N, x, None, brk, t_q, bw_q, dtheta=e, threads=2)
    ref, A, _ = oracle_mod.breakage_rhs(N, x, t_q, bw_q, **dict(brk, S0=1.0))
    assert_parity(dR, ref, A, tol=1e-13, what=f"dR/dS0 sel {selection}")


@pytest.mark.parametrize("kernel,selection", [(5, 2), (0, 3), (2, 1)])
def test_cs_population_derivative_central_difference(oracle_mod, kernel, selection):
    """R is quadratic in N, so (R(N + d) - R(N - d)) / 2 = dR/dN . d exactly."""
    x, N, agg, brk, t_q, bw_q = _case(kernel=kernel, selection=selection)
    # |d| < N keeps N +- d positive, inside the breakage guard N > 2.2e-308
    d = np.random.default_rng(kernel).uniform(-0.5, 0.5, size=len(N)) * N
    _, dR = oracle_mod.cs_rhs(N, x, agg, brk, t_q, bw_q, dN=d, threads=2)
    rp, Ap = _real_rhs(oracle_mod, N + d, x, agg, brk, t_q, bw_q)
    rm, Am = _real_rhs(oracle_mod, N - d, x, agg, brk, t_q, bw_q)
    fd = 0.5 * (rp - rm)
    scale = Ap + Am + np.abs(dR)
    nz = scale > 0
    err = np.abs(dR - fd)[nz] / scale[nz]
    assert err.max() <= 1e-13, err.max()


def test_cs_alpha_derivative_central_difference(oracle_mod):
    x, N, agg, brk, t_q, bw_q = _case(selection=2)
    e = np.zeros(5)
    e[4] = 1.0
    _, dR = oracle_mod.cs_rhs(N, x, None, brk, t_q, bw_q, dtheta=e, threads=2)
    a0 = brk["alpha"]
    h = 1e-5 * a0
    rp, Ap, _ = oracle_mod.breakage_rhs(N, x, t_q, bw_q, **dict(brk, alpha=a0 + h))
    rm, Am, _ = oracle_mod.breakage_rhs(N, x, t_q, bw_q, **dict(brk, alpha=a0 - h))
    fd = (rp - rm) / (2 * h)
    scale = np.abs(dR).max()
    assert scale > 0
    assert np.max(np.abs(dR - fd)) <= 1e-8 * scale


def test_cs_loss_against_real_rk4_and_central_differences(oracle_mod):
    """cs_loss: its real part is the real oracle's RK4 followed by the log
    misfit, and each derivative matches central differences of that."""
    from pbe_hip import problems as P
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(96)
    # on this short grid (x <= 2e-17) shear needs a large coefficient to move
    # the loss as much as the other three parameters do
    agg = dict(agg, beta_sh=1e9)
    dt, steps = 0.5, 40
    th_true = np.array([agg["beta_br"] * 1.7, agg["beta_sh"] * 0.6, brk["S0"] * 1.4,
                        brk["alpha"] * 1.3])
    a_t, b_t = P.with_params(agg, brk, th_true)
    T = oracle_mod.rk4(N0, x, dt, steps, agg=a_t, brk=b_t, t_q=t_q, bw_q=bw_q)
    eps = 1e-30 * N0.sum()
    logT = np.log(np.maximum(T, eps))
    mask = (T > 1e-8 * T.max()).astype(float)

    def real_loss(th, N=N0):
        a, b = P.with_params(agg, brk, th)
        Nf = oracle_mod.rk4(N, x, dt, steps, agg=a, brk=b, t_q=t_q, bw_q=bw_q)
        return P.log_misfit(Nf, logT, mask, eps)[0]

    th0 = np.array([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]])
    L0 = real_loss(th0)
    for p, slot in enumerate((1, 2, 3, 4)):
        e = np.zeros(5)
        e[slot] = 1.0
        Lc, dL = oracle_mod.cs_loss(N0, x, dt, steps, logT, mask, eps, agg, brk, t_q, bw_q,
                                    dtheta=e, threads=4)
        assert abs(Lc - L0) <= 1e-11 * abs(L0)
        h = 1e-5 * th0[p]
        tp, tm = th0.copy(), th0.copy()
        tp[p] += h
        tm[p] -= h
        fd = (real_loss(tp) - real_loss(tm)) / (2 * h)
        assert abs(dL * th0[p]) >= 0.1 * L0   # each parameter moves the loss
        assert abs(dL - fd) <= 1e-7 * abs(dL), (P.FIT_PARAMS[p], dL, fd)
    dN0 = np.random.default_rng(7).normal(size=len(N0)) * N0
    _, dL = oracle_mod.cs_loss(N0, x, dt, steps, logT, mask, eps, agg, brk, t_q, bw_q,
                               dN0=dN0, threads=4)
    h = 1e-5
    fd = (real_loss(th0, N0 + h * dN0) - real_loss(th0, N0 - h * dN0)) / (2 * h)
    assert abs(dL - fd) <= 1e-7 * abs(dL), (dL, fd)
