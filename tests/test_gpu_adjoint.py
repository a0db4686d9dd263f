"""Hand-written VJP and checkpointed discrete RK4 adjoint, checked against
complex-step derivatives of the CPU oracle (machine-precision derivatives of
the same discrete map) and central finite differences of GPU forward solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_problem(n=160, kernel=5, selection=2):
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_adjoint(n)
    agg = dict(agg, kernel=kernel, beta0=1e-17 if kernel < 3 else 0.0)
    if kernel == 1:
        agg["beta0"] = 1e-17 / x[n // 2]
    if kernel == 2:
        agg["beta0"] = 1e-17 / x[n // 2] ** 2
    brk = dict(brk, selection=selection, v_min=x[n // 2])
    return x, N, agg, brk, t_q, bw_q


@pytest.mark.parametrize("kernel,selection,daughter", [
    (5, 2, "symmetric"), (0, 0, "uniform"), (1, 1, "erosion"), (2, 3, "powerlaw"),
    (3, 2, "uniform"), (4, 1, "symmetric")])
def test_vjp_vs_complex_step(pbe, oracle_mod, kernel, selection, daughter):
    from pbe_hip import problems as P
    x, N, agg, brk, _, _ = _small_problem(200, kernel, selection)
    t_q, bw_q = P.daughter_table(daughter, 8, q=2.0, eps=0.05)
    rng = np.random.default_rng(kernel)
    lam = rng.normal(size=len(N))
    gN, gth = pbe.rhs_vjp(N, x, lam, agg, brk, t_q, bw_q)
    for trial in range(3):
        dN = rng.normal(size=len(N)) * (N + 1e-3 * N.max())
        _, dR = oracle_mod.cs_rhs(N, x, agg, brk, t_q, bw_q, dN=dN)
        ref = float(lam @ dR)
        got = float(gN @ dN)
        scale = float(np.abs(lam) @ np.abs(dR)) + abs(ref)
        assert abs(got - ref) <= 1e-11 * scale, (trial, got, ref)
    for p in range(5):
        e = np.zeros(5)
        e[p] = 1.0
        _, dR = oracle_mod.cs_rhs(N, x, agg, brk, t_q, bw_q, dtheta=e)
        ref = float(lam @ dR)
        scale = float(np.abs(lam) @ np.abs(dR))
        assert abs(gth[p] - ref) <= 1e-11 * scale + 1e-300, (p, gth[p], ref)


def _target(x, N0, agg, brk, t_q, bw_q, dt, steps, seed):
    """Target PSD from perturbed true parameters (each x U[0.5, 2])."""
    from pbe_hip import hip, problems as P
    import pbe_hip as pbe
    rng = np.random.default_rng(seed)
    th0 = np.array([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]])
    th_true = th0 * rng.uniform(0.5, 2.0, 4)
    a, b = P.with_params(agg, brk, th_true)
    it = pbe.Integrator(x, agg=a, brk=b, t_q=t_q, bw_q=bw_q)
    d = hip.DeviceArray.from_host(N0)
    hip.check(it.step(d, dt, steps))
    hip.synchronize()
    T = d.download()
    eps = 1e-30 * N0.sum()
    mask = (T > 1e-8 * T.max()).astype(float)
    return np.log(np.maximum(T, eps)), mask, eps


def _adjoint_grad(pbe, x, N0, agg, brk, t_q, bw_q, dt, steps, every, logT, mask, eps):
    from pbe_hip import hip, problems as P
    adj = pbe.Adjoint(x, dt, steps, every, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(adj.forward(dN))
    hip.synchronize()
    Nf = dN.download()
    L, g = P.log_misfit(Nf, logT, mask, eps)
    dg, gN0, gth = hip.DeviceArray.from_host(g), hip.DeviceArray(len(N0)), hip.DeviceArray(5)
    hip.check(adj.backward(dg, gN0, gth))
    hip.synchronize()
    return L, gN0.download(), gth.download()


@pytest.mark.parametrize("every", [7, 0])
def test_adjoint_rk4_vs_complex_step(pbe, oracle_mod, every):
    """Small instance of the fit problem: every gradient component of the
    discrete RK4 map against complex-step on the CPU, to 1e-8 relative.
    A checkpoint interval that does not divide the step count exercises the
    ragged last segment; every = 0 keeps all stage states (no recompute)."""
    from pbe_hip import problems as P
    x, N0, agg, brk, t_q, bw_q = _small_problem(128)
    dt, steps = 0.5, 60
    logT, mask, eps = _target(x, N0, agg, brk, t_q, bw_q, dt, steps, seed=1)
    L, gN0, gth = _adjoint_grad(pbe, x, N0, agg, brk, t_q, bw_q, dt, steps, every, logT,
                                mask, eps)
    for p, slot in ((0, 1), (1, 2), (2, 3), (3, 4)):
        e = np.zeros(5)
        e[slot] = 1.0
        Lc, dL = oracle_mod.cs_loss(N0, x, dt, steps, logT, mask, eps, agg, brk, t_q, bw_q,
                                    dtheta=e, threads=4)
        assert abs(Lc - L) <= 1e-10 * abs(L)
        assert abs(gth[slot] - dL) <= 1e-8 * abs(dL), (P.FIT_PARAMS[p], gth[slot], dL)
    rng = np.random.default_rng(3)
    dN0 = rng.normal(size=len(N0)) * N0
    _, dL = oracle_mod.cs_loss(N0, x, dt, steps, logT, mask, eps, agg, brk, t_q, bw_q,
                               dN0=dN0, threads=4)
    assert abs(gN0 @ dN0 - dL) <= 1e-8 * abs(dL)


def test_adjoint_config5_vs_central_fd(pbe):
    """n = 1024, r = 1.03, Brownian+shear + power-law breakage, 2000 RK4 steps of
    0.5 s, checkpoints every 50: parameter gradient against central finite
    differences (h = 1e-6 relative) of GPU forward solves."""
    from pbe_hip import hip, problems as P
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(1024)
    dt, steps, every = 0.5, 2000, 50
    logT, mask, eps = _target(x, N0, agg, brk, t_q, bw_q, dt, steps, seed=2)
    L, gN0, gth = _adjoint_grad(pbe, x, N0, agg, brk, t_q, bw_q, dt, steps, every, logT,
                                mask, eps)
    th0 = np.array([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]])

    def loss(th):
        a, b = P.with_params(agg, brk, th)
        it = pbe.Integrator(x, agg=a, brk=b, t_q=t_q, bw_q=bw_q)
        d = hip.DeviceArray.from_host(N0)
        hip.check(it.step(d, dt, steps))
        hip.synchronize()
        return P.log_misfit(d.download(), logT, mask, eps)[0]

    for p, slot in enumerate((1, 2, 3, 4)):
        h = 1e-6 * th0[p]
        tp, tm = th0.copy(), th0.copy()
        tp[p] += h
        tm[p] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        print(P.FIT_PARAMS[p], gth[slot], fd)
        assert abs(gth[slot] - fd) <= 1e-5 * abs(fd) + 1e-9 * abs(L) / abs(th0[p]), p


@pytest.mark.parametrize("every", [50, 0])
def test_adjoint_config5_vs_complex_step_fixture(pbe, every):
    """Config 5 at full size (n = 1024, r = 1.03, 2000 RK4 steps of 0.5 s):
    the GPU adjoint, with checkpoints every 50 steps and with every stage
    kept, against the complex-step gradient of the same discrete map computed
    by the CPU oracle (tests/golden/adjoint_config5_cs.json, written by
    tests/golden/make_adjoint_golden.py): the four parameter derivatives and
    one direction of dL/dN0, each to 1e-8 relative."""
    import json
    import os
    from pbe_hip import problems as P
    with open(os.path.join(os.path.dirname(__file__), "golden", "adjoint_config5_cs.json")) as f:
        fx = json.load(f)
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(fx["n"])
    assert np.allclose([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]],
                       fx["theta0"], rtol=0, atol=0)
    logT, mask = np.array(fx["logT"]), np.array(fx["mask"])
    L, gN0, gth = _adjoint_grad(pbe, x, N0, agg, brk, t_q, bw_q, fx["dt"], fx["steps"], every,
                                logT, mask, fx["eps"])
    assert abs(L - fx["L"]) <= 1e-10 * abs(fx["L"]), (L, fx["L"])
    for p, slot in enumerate((1, 2, 3, 4)):
        ref = fx["grad"][P.FIT_PARAMS[p]]
        print(P.FIT_PARAMS[p], gth[slot], ref)
        assert abs(gth[slot] - ref) <= 1e-8 * abs(ref), (P.FIT_PARAMS[p], gth[slot], ref)
    g_dir = float(gN0 @ np.array(fx["dN0"]))
    assert abs(g_dir - fx["grad_dN0"]) <= 1e-8 * abs(fx["grad_dN0"]), (g_dir, fx["grad_dN0"])


def _batch_small(pbe, count, seed, n_max):
    from pbe_hip import problems as P
    inst, N, x, t_q, bw_q = P.config_batched_fit(count, seed=seed, n_min=64, n_max=n_max)
    plan = pbe.BatchPlan(inst, kernel_type=5, selection=2, n_quad=len(t_q))
    return inst, N, x, t_q, bw_q, plan


def test_batched_vjp_matches_single(pbe):
    """The batched VJP per instance equals the single-instance VJP (the two
    cut the column range into different chunks, so the order of the partial
    sums differs: 1e-13 of the instance's largest entry); each is bitwise
    repeatable; the joint-scale gradient is sum_i coef_i dL/dcoef_i."""
    from pbe_hip import problems as P
    inst, N, x, t_q, bw_q, plan = _batch_small(pbe, 12, 21, 400)
    lam = np.random.default_rng(2).normal(size=len(N))
    gN, gth = pbe.batch_rhs_vjp(plan, N, x, t_q, bw_q, lam)
    gN2, gth2 = pbe.batch_rhs_vjp(plan, N, x, t_q, bw_q, lam)
    assert np.array_equal(gN, gN2) and np.array_equal(gth, gth2)
    expect = np.zeros(5)
    for s in inst:
        lo, hi = s["offset"], s["offset"] + s["n"]
        agg, brk = P.instance_params(s)
        g1, t1 = pbe.rhs_vjp(N[lo:hi], x[lo:hi], lam[lo:hi], agg, brk, t_q, bw_q)
        g1b, t1b = pbe.rhs_vjp(N[lo:hi], x[lo:hi], lam[lo:hi], agg, brk, t_q, bw_q)
        assert np.array_equal(g1, g1b) and np.array_equal(t1, t1b)
        assert np.max(np.abs(gN[lo:hi] - g1)) <= 1e-13 * np.max(np.abs(g1)), s["n"]
        expect += t1 * np.array([s["beta0"], s["beta_br"], s["beta_sh"], s["S0"], 1.0])
    assert np.allclose(gth, expect, rtol=1e-11, atol=0)


@pytest.mark.parametrize("steps,every", [(1, 50), (5, 1), (3, 0)])
def test_adjoint_short_runs(pbe, oracle_mod, steps, every):
    """Edge cases of the reverse sweep: a single step (its last stage also
    starts the step that never comes), a checkpoint interval beyond the step
    count (one segment) and one step per segment."""
    from pbe_hip import problems as P
    x, N0, agg, brk, t_q, bw_q = _small_problem(96)
    dt = 0.5
    logT, mask, eps = _target(x, N0, agg, brk, t_q, bw_q, dt, steps, seed=2)
    L, gN0, gth = _adjoint_grad(pbe, x, N0, agg, brk, t_q, bw_q, dt, steps, every, logT,
                                mask, eps)
    for p, slot in ((0, 1), (1, 2), (2, 3), (3, 4)):
        e = np.zeros(5)
        e[slot] = 1.0
        Lc, dL = oracle_mod.cs_loss(N0, x, dt, steps, logT, mask, eps, agg, brk, t_q, bw_q,
                                    dtheta=e, threads=4)
        assert abs(Lc - L) <= 1e-10 * abs(L)
        assert abs(gth[slot] - dL) <= 1e-8 * abs(dL), (P.FIT_PARAMS[p], steps, every)
    rng = np.random.default_rng(5)
    dN0 = rng.normal(size=len(N0)) * N0
    _, dL = oracle_mod.cs_loss(N0, x, dt, steps, logT, mask, eps, agg, brk, t_q, bw_q,
                               dN0=dN0, threads=4)
    assert abs(gN0 @ dN0 - dL) <= 1e-8 * abs(dL), (steps, every)


@pytest.mark.parametrize("every", [8, 0])
def test_batched_adjoint_vs_complex_step(pbe, oracle_mod, every):
    """Joint fit of a small batch: d/d ln(scale) of the summed log misfit
    against complex-step of every instance on the CPU."""
    from pbe_hip import hip, problems as P
    inst, N0, x, t_q, bw_q, plan = _batch_small(pbe, 5, 33, 160)
    dt, steps = 0.5, 30
    eps = 1e-30 * 1e12
    # target: one forward with the scales perturbed (all instances alike)
    logT = np.zeros(len(N0))
    mask = np.zeros(len(N0))
    scale = np.array([1.0, 1.6, 0.7, 1.3])
    for s in inst:
        lo, hi = s["offset"], s["offset"] + s["n"]
        agg, brk = P.instance_params(s)
        a, b = P.with_params(agg, brk, [agg["beta_br"] * scale[0], agg["beta_sh"] * scale[1],
                                        brk["S0"] * scale[2], brk["alpha"] * scale[3]])
        T = oracle_mod.rk4(N0[lo:hi], x[lo:hi], dt, steps, agg=a, brk=b, t_q=t_q, bw_q=bw_q)
        logT[lo:hi] = np.log(np.maximum(T, eps))
        mask[lo:hi] = T > 1e-8 * T.max()
    adj = pbe.BatchAdjoint(plan, x, t_q, bw_q, dt, steps, every)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(adj.forward(dN))
    hip.synchronize()
    L, g = P.log_misfit(dN.download(), logT, mask, eps)
    dg, gN0, gth = hip.DeviceArray.from_host(g), hip.DeviceArray(len(N0)), hip.DeviceArray(5)
    hip.check(adj.backward(dg, gN0, gth))
    hip.synchronize()
    gth = gth.download()
    expect = np.zeros(5)
    for s in inst:
        lo, hi = s["offset"], s["offset"] + s["n"]
        agg, brk = P.instance_params(s)
        coef = [0.0, s["beta_br"], s["beta_sh"], s["S0"], 1.0]
        for slot in (1, 2, 3, 4):
            e = np.zeros(5)
            e[slot] = coef[slot]
            _, dL = oracle_mod.cs_loss(N0[lo:hi], x[lo:hi], dt, steps, logT[lo:hi],
                                       mask[lo:hi], eps, agg, brk, t_q, bw_q, dtheta=e)
            expect[slot] += dL
    for slot in (1, 2, 3, 4):
        assert abs(gth[slot] - expect[slot]) <= 1e-8 * abs(expect[slot]), (slot, gth, expect)


def test_vjp_breakage_only_vs_complex_step(pbe, oracle_mod):
    from pbe_hip import problems as P
    x, N, _, brk, _, _ = _small_problem(150, 5, 2)
    t_q, bw_q = P.daughter_power_law(2.0, 8)
    lam = np.random.default_rng(4).normal(size=len(N))
    gN, gth = pbe.rhs_vjp(N, x, lam, None, brk, t_q, bw_q)
    dN = np.random.default_rng(5).normal(size=len(N)) * N
    _, dR = oracle_mod.cs_rhs(N, x, None, brk, t_q, bw_q, dN=dN)
    assert abs(gN @ dN - lam @ dR) <= 1e-11 * (np.abs(lam) @ np.abs(dR))
    for slot in (3, 4):
        e = np.zeros(5)
        e[slot] = 1.0
        _, dR = oracle_mod.cs_rhs(N, x, None, brk, t_q, bw_q, dtheta=e)
        assert abs(gth[slot] - lam @ dR) <= 1e-11 * (np.abs(lam) @ np.abs(dR))
