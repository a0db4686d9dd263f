"""Device time integration (pbe_integrate) and its hipGraph against the
oracle's RK4 (the reference harness arithmetic,
examples/common/pbe_examples_utils.hpp:416-457) and the Smoluchowski closed
form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_smoluchowski_rk4_closed_form(pbe, oracle_mod):
    """Constant kernel, n = 256, x_i = 1e-18 2^(i/4), dt = 1 s, 1000 steps:
    M0(t) = M0(0)/(1 + beta0 M0(0) t/2) to RK4 accuracy (4.2e-7 expected)."""
    from pbe_hip import hip, problems as P
    x, N0, agg = P.config_smoluchowski(256)
    it = pbe.Integrator(x, agg=agg)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, 1.0, 1000))
    hip.synchronize()
    N = dN.download()
    M00 = N0.sum()
    exact = M00 / (1.0 + 1e-12 * M00 * 1000.0 / 2.0)
    assert abs(N.sum() - exact) / exact <= 1e-6
    ref = oracle_mod.rk4(N0, x, 1.0, 1000, agg=agg)
    scale = np.abs(ref).max()
    err = np.abs(N - ref).max() / scale
    print("rk4 1000 steps, max |gpu-oracle|/max|N| =", err)
    assert err <= 1e-11


@pytest.mark.parametrize("case", ["sum", "product", "breakage"])
def test_rk4_moment_laws(pbe, case):
    """Device RK4 against the discrete moment laws: the sum kernel's M0
    decays by RK4's own polynomial, the product kernel's and linear-selection
    breakage's M0 is linear in t, M1 is conserved (Golovin, product-kernel
    and Ziff-McGrady moment behaviour of the fixed-pivot system)."""
    from pbe_hip import hip
    import moment_cases as MC
    x, N0, kw, M0_expected, M1 = MC.CASES[case]()
    it = pbe.Integrator(x, **kw)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, MC.DT, MC.STEPS))
    hip.synchronize()
    MC.check(x, dN.download(), M0_expected, M1, tol_m0=1e-11)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_reference_examples(pbe, k):
    """The reference's five example programs run on the device integrator
    (explicit Euler or RK4, as each example does): the documented errors
    against the analytic moments are reproduced (docs/examples.md)."""
    from pbe_hip import hip
    import reference_examples as RE
    c = RE.example(k)
    method = pbe.METHOD_RK4 if c["method"] == "rk4" else pbe.METHOD_EULER
    it = pbe.Integrator(c["x"], method=method, **c["kw"])
    dN = hip.DeviceArray.from_host(c["N0"])
    hip.check(it.step(dN, c["dt"], c["steps"]))
    hip.synchronize()
    print(k, RE.assert_matches(c, dN.download()))


def test_flocculation_rk4_vs_oracle(pbe, oracle_mod):
    from pbe_hip import hip, problems as P
    x, N0, agg, brk, _, _ = P.config_adjoint(512)
    t_q, bw_q = P.daughter_uniform(8)
    it = pbe.Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, 0.5, 40))
    hip.synchronize()
    N = dN.download()
    ref = oracle_mod.rk4(N0, x, 0.5, 40, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    assert np.all(np.isfinite(N))
    err = np.abs(N - ref).max() / np.abs(ref).max()
    print("agg+brk rk4 40 steps:", err)
    assert err <= 1e-11


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("selection", [0, 1, 2, 3])
def test_fused_rhs_every_kernel_and_selection(pbe, oracle_mod, kernel, selection):
    """Small problems evaluate R_agg + R_brk in one launch (aggregation tile
    items and breakage parent blocks in one grid): every collision kernel x
    selection function under the integrator against the oracle's RK4 and
    Euler, on a geometric (n = 200) and a jittered (n = 333, binary-search
    brackets) grid, with an aggregation-only and a breakage-only problem
    (separate launches) beside it."""
    from pbe_hip import hip, problems as P
    import test_gpu_adjoint as TA
    x, N0, agg, brk, _, _ = TA._small_problem(200, kernel, selection)
    t_q, bw_q = P.daughter_uniform(8)
    xj = P.jittered_grid(333, 1e-18, 1.03, 0.2, seed=kernel * 4 + selection)
    Nj = P.lognormal_ic(xj, xj[200], 0.8, 1e4, noise=0.05, seed=3)
    aj = dict(agg, log_x0=0.0, inv_log_r=0.0)
    bj = dict(brk, v_ref=float(xj[166]), v_min=float(xj[166]))
    cases = [(x, N0, agg, brk), (xj, Nj, aj, bj), (x, N0, agg, None), (x, N0, None, brk)]
    for ci, (xx, NN, a, b) in enumerate(cases):
        kw = dict(agg=a, brk=b, t_q=t_q if b else None, bw_q=bw_q if b else None)
        for method, steps in ((pbe.METHOD_RK4, 6), (pbe.METHOD_EULER, 3)):
            it = pbe.Integrator(xx, method=method, **kw)
            dN = hip.DeviceArray.from_host(NN)
            hip.check(it.step(dN, 0.5, steps))
            hip.synchronize()
            got = dN.download()
            if method == pbe.METHOD_RK4:
                ref = oracle_mod.rk4(NN, xx, 0.5, steps, **kw)
            else:
                ref = NN.copy()
                for _ in range(steps):
                    r = np.zeros_like(ref)
                    if a:
                        r += oracle_mod.aggregation_rhs(ref, xx, **a)[0]
                    if b:
                        r += oracle_mod.breakage_rhs(ref, xx, t_q, bw_q, **b)[0]
                    ref = ref + 0.5 * r
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err <= 1e-11, (ci, method, err)


@pytest.mark.parametrize("n", [3072, 4096, 6144])
def test_fused_rhs_mid_sizes(pbe, oracle_mod, n):
    """Mid-size flocculation problems (Brownian+shear + threshold breakage,
    r = 1.02) whose aggregation the tuned table runs as one tile launch of
    128- or 256-column tiles also evaluate R_agg + R_brk in one launch under
    the integrator: one Euler step per entry against N + dt (R_agg + R_brk)
    of the oracle, and 2 RK4 steps against the oracle's RK4, with dt a
    twentieth of max N / max |R| (the throughput workload's rates would blow
    up a step of seconds)."""
    from pbe_hip import hip, problems as P
    x, N0, agg, brk, t_q, bw_q = P.config_flocculation(n)
    kw = dict(agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    ra, Aa, _ = oracle_mod.aggregation_rhs(N0, x, threads=16, **agg)
    rb, Ab, _ = oracle_mod.breakage_rhs(N0, x, t_q, bw_q, **brk)
    dt = 0.05 * np.abs(N0).max() / np.abs(ra + rb).max()
    print(n, pbe.aggregation_default_options(pbe.agg_params(n, agg)), dt)

    it = pbe.Integrator(x, method=pbe.METHOD_EULER, **kw)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, dt, 1))
    hip.synchronize()
    expect = N0 + dt * (ra + rb)
    tol = 1e-12 * dt * (Aa + Ab) + 1e-15 * np.abs(N0)
    assert np.all(np.abs(dN.download() - expect) <= tol + 1e-300)

    it = pbe.Integrator(x, method=pbe.METHOD_RK4, **kw)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, dt, 2))
    hip.synchronize()
    got = dN.download()
    ref = oracle_mod.rk4(N0, x, dt, 2, threads=16, **kw)
    assert np.all(np.isfinite(ref))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("rk4 2 steps:", err)
    assert err <= 1e-11, err


def test_euler_vs_oracle_one_step(pbe, oracle_mod):
    from pbe_hip import hip, problems as P
    x, N0, agg, brk, t_q, bw_q = P.config_flocculation(300)
    it = pbe.Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q, method=pbe.METHOD_EULER)
    dN = hip.DeviceArray.from_host(N0)
    hip.check(it.step(dN, 0.25, 1))
    hip.synchronize()
    ra, Aa, _ = oracle_mod.aggregation_rhs(N0, x, **agg)
    rb, Ab, _ = oracle_mod.breakage_rhs(N0, x, t_q, bw_q, **brk)
    expect = N0 + 0.25 * (ra + rb)
    tol = 1e-12 * 0.25 * (Aa + Ab) + 1e-15 * np.abs(N0)
    assert np.all(np.abs(dN.download() - expect) <= tol + 1e-300)


def test_graph_replays_match_eager(pbe):
    """The captured RK4 graph, replayed several times with eager work and
    synchronisations in between, tracks the eager integration after every
    replay (no memset node is recorded)."""
    from pbe_hip import hip, problems as P
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(1024)
    it_g = pbe.Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    it_e = pbe.Integrator(x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    dNg = hip.DeviceArray.from_host(N0)
    dNe = hip.DeviceArray.from_host(N0)
    g = it_g.graph(dNg, 0.5, 5)
    s = hip.Stream()
    for rep in range(4):
        hip.check(g.launch(s))
        s.synchronize()
        hip.check(it_e.step(dNe, 0.5, 5))
        hip.synchronize()
        a, b = dNg.download(), dNe.download()
        assert np.all(np.isfinite(a))
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err <= 1e-12, (rep, err)
        assert not np.array_equal(a, N0)
    g.destroy()


def test_integrator_argument_errors(pbe):
    from pbe_hip import hip
    L = pbe.library()
    p = pbe.AggregationParams(n=10)
    assert L.pbe_integrate(1, 1, None, None, None, None, 1.0, 1, 0, 1, 10 ** 6, None) == 1
    import ctypes
    assert L.pbe_integrate(1, None, None, None, ctypes.byref(p), None, 1.0, 1, 0, 1, 10 ** 6,
                           None) == 17
    assert L.pbe_integrate(1, 1, None, None, ctypes.byref(p), None, 1.0, 1, 0, 1, 8,
                           None) == 1
