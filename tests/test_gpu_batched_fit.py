"""Joint parameter fit over a batch of 256 instances (n log-uniform in
[128, 1024], each with its own shear rate and initial PSD) through the
batched discrete RK4 adjoint: the gradient against FP64 complex-step of the
CPU oracle's RK4 on a few instances, and an L-BFGS fit recovering common
parameter factors from a target solved at known ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 0.5


def _fit_problem(steps, every=50):
    from pbe_hip import problems as P
    from pbe_hip.fit import BatchedFit
    inst, N0, x, t_q, bw_q = P.config_batched_fit(256, seed=256)
    return BatchedFit(inst, N0, x, t_q, bw_q, DT, steps, every=every)


def _theta_true(seed):
    rng = np.random.default_rng(seed)
    return np.array([*np.log(rng.uniform(0.5, 2.0, 3)), (1.0 / 3.0) * rng.uniform(0.5, 2.0)])


def test_batched_fit_gradient_vs_complex_step(pbe, oracle_mod):
    """The loss restricted (by its mask) to three of the 256 instances: its
    batched-adjoint gradient is the sum of those instances' complex-step
    gradients (d/d ln s = coef * d/d coef for the three scaled
    coefficients, d/d alpha directly), to 1e-8 relative, with checkpoints
    every 50 steps and with every stage state kept."""
    from pbe_hip import problems as P
    steps = 100
    F = _fit_problem(steps)
    F.set_target(_theta_true(3))
    order = np.argsort([d["n"] for d in F.base])
    picks = [int(order[0]), int(order[40]), int(order[90])]
    mask = np.zeros_like(F.mask)
    for i in picks:
        d = F.base[i]
        lo, hi = d["offset"], d["offset"] + d["n"]
        mask[lo:hi] = F.mask[lo:hi]
    L, g = F.loss_and_grad(F.theta0, mask=mask)
    expect = np.zeros(4)
    Ls = 0.0
    for i in picks:
        d = F.base[i]
        lo, hi = d["offset"], d["offset"] + d["n"]
        agg, brk = P.instance_params(d)
        coef = (d["beta_br"], d["beta_sh"], d["S0"], 1.0)
        for p, slot in enumerate((1, 2, 3, 4)):
            e = np.zeros(5)
            e[slot] = coef[p]
            Lc, dL = oracle_mod.cs_loss(F.N0[lo:hi], F.x[lo:hi], DT, steps, F.logT[lo:hi],
                                        mask[lo:hi], F.eps, agg, brk, F.t_q, F.bw_q,
                                        dtheta=e, threads=4)
            expect[p] += dL
        Ls += Lc
    assert abs(L - Ls) <= 1e-10 * abs(Ls), (L, Ls)
    for p in range(4):
        assert abs(g[p] - expect[p]) <= 1e-8 * abs(expect[p]), (p, g, expect)
    F0 = _fit_problem(steps, every=0)
    F0.logT, F0.mask = F.logT, F.mask
    L0, g0 = F0.loss_and_grad(F0.theta0, mask=mask)
    assert abs(L0 - Ls) <= 1e-10 * abs(Ls), (L0, Ls)
    for p in range(4):
        assert abs(g0[p] - expect[p]) <= 1e-8 * abs(expect[p]), ("store all", p, g0, expect)


def test_batched_joint_fit_converges(pbe):
    """L-BFGS from all factors 1 to the factors the target was solved at
    (each of beta_br, beta_sh, S0 x U[0.5, 2], alpha x U[0.5, 2]), 200 RK4
    steps of 0.5 s over all 256 instances."""
    F = _fit_problem(200)
    th_true = _theta_true(11)
    F.set_target(th_true)
    th, hist, res = F.fit(max_evals=60)
    print("evaluations", len(hist), "loss", hist[0], "->", hist[-1], "theta", th, th_true)
    assert hist[-1] <= 1e-12 * hist[0]
    assert np.max(np.abs(th - th_true)) <= 1e-6, (th, th_true)
