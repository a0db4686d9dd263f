"""Batched form: one persistent launch over many instances of uneven n,
against the oracle per instance, against a loop of the plain per-instance
calls, and under RK4 / hipGraph."""
import numpy as np
import pytest

from parity import assert_parity

pytestmark = pytest.mark.gpu


def _setup(pbe, count, seed, n_max=2048, fit=False):
    from pbe_hip import hip, problems as P
    if fit:
        inst, N, x, t_q, bw_q = P.config_batched_fit(count, seed=seed, n_max=n_max)
    else:
        inst, N, x, t_q, bw_q = P.config_batched(count, seed=seed, n_max=n_max)
    plan = pbe.BatchPlan(inst, kernel_type=5, selection=2, n_quad=len(t_q))
    assert plan.size == len(N)
    d = dict(N=hip.DeviceArray.from_host(N), x=hip.DeviceArray.from_host(x),
             t=hip.DeviceArray.from_host(t_q), b=hip.DeviceArray.from_host(bw_q))
    return inst, N, x, t_q, bw_q, plan, d


def test_batched_rhs_vs_oracle(pbe, oracle_mod):
    from pbe_hip import hip, problems as P
    inst, N, x, t_q, bw_q, plan, d = _setup(pbe, 96, seed=5)
    rhs = hip.DeviceArray(len(N))
    rhs.zero()
    hip.check(plan.launch(d["N"], d["x"], d["t"], d["b"], rhs))
    hip.synchronize()
    got = rhs.download()
    worst = 0.0
    for s in inst:
        lo, hi = s["offset"], s["offset"] + s["n"]
        agg, brk = P.instance_params(s)
        ra, Aa, _ = oracle_mod.aggregation_rhs(N[lo:hi], x[lo:hi], **agg)
        rb, Ab, _ = oracle_mod.breakage_rhs(N[lo:hi], x[lo:hi], t_q, bw_q, **brk)
        worst = max(worst, assert_parity(got[lo:hi], ra + rb, Aa + Ab,
                                         what=f"instance n={s['n']}"))
    print("batched worst entry error", worst)


def test_batched_matches_plain_loop(pbe):
    from pbe_hip import hip, problems as P
    inst, N, x, t_q, bw_q, plan, d = _setup(pbe, 300, seed=9)
    rb = hip.DeviceArray(len(N))
    rb.zero()
    hip.check(plan.launch(d["N"], d["x"], d["t"], d["b"], rb))
    rl = hip.DeviceArray(len(N))
    rl.zero()
    for s in inst:
        agg, brk = P.instance_params(s)
        off = s["offset"] * 8
        hip.check(pbe.launch_aggregation_rhs(d["N"].value + off, d["x"].value + off,
                                             rl.value + off, pbe.agg_params(s["n"], agg)))
        hip.check(pbe.launch_breakage_rhs(d["N"].value + off, d["x"].value + off, d["t"],
                                          d["b"], rl.value + off,
                                          pbe.brk_params(s["n"], len(t_q), brk)))
    hip.synchronize()
    a, b = rb.download(), rl.download()
    scale = np.maximum(np.abs(a), np.abs(b))
    # same arithmetic per entry, different atomic order: agreement to rounding
    assert np.all(np.abs(a - b) <= 1e-12 * np.max(scale) + 1e-12 * scale)


def test_batched_rk4_and_graph(pbe):
    from pbe_hip import hip, problems as P
    inst, N, x, t_q, bw_q, plan, d = _setup(pbe, 64, seed=13, n_max=1024, fit=True)
    L = pbe.library()
    wsb = L.pbe_integrator_workspace_bytes(len(N), pbe.METHOD_RK4)
    ws = hip.DeviceArray(wsb // 8)
    hip.check(plan.integrate(d["N"], d["x"], d["t"], d["b"], 0.5, 10, ws, wsb))
    hip.synchronize()
    got = d["N"].download()
    # the same 10 steps per instance through the plain integrator
    for s in inst[:12]:
        lo, hi = s["offset"], s["offset"] + s["n"]
        agg, brk = P.instance_params(s)
        it = pbe.Integrator(x[lo:hi], agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
        dn = hip.DeviceArray.from_host(N[lo:hi])
        hip.check(it.step(dn, 0.5, 10))
        hip.synchronize()
        ref = dn.download()
        assert np.abs(got[lo:hi] - ref).max() <= 1e-12 * np.abs(ref).max()
    # graph replays track the eager batched integration
    Ng = hip.DeviceArray.from_host(N)
    wsg = hip.DeviceArray(wsb // 8)
    g = plan.graph(Ng, d["x"], d["t"], d["b"], 0.5, 5, wsg, wsb)
    Ne = hip.DeviceArray.from_host(N)
    for rep in range(3):
        hip.check(g.launch())
        hip.synchronize()
        hip.check(plan.integrate(Ne, d["x"], d["t"], d["b"], 0.5, 5, ws, wsb))
        hip.synchronize()
        a, b = Ng.download(), Ne.download()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), rep
    g.destroy()


@pytest.mark.parametrize("tile_u", [1, 2, 4, 8])
@pytest.mark.parametrize("kernel", [5, 2])
def test_batched_every_tile_width_vs_oracle(pbe, oracle_mod, tile_u, kernel):
    """Every forced tile width of the batched aggregation (64 tile_u wide
    pair tiles, one launch per width) against the oracle per instance, on
    geometric instances and on instances without the geometric lookup
    (inv_log_r = 0, jittered grids), n from 2 to 1100."""
    from pbe_hip import hip, problems as P
    rng = np.random.default_rng(100 * tile_u + kernel)
    inst, xs, Ns = [], [], []
    off = 0
    for i, n in enumerate([2, 3, 63, 64, 65, 129, 255, 256, 257, 511, 700, 1100]):
        general = i % 3 == 2
        x = (P.jittered_grid(n, 1e-18, 1.03, 0.3, seed=i) if general
             else P.geometric_grid(n, 1e-18, 1.03 if i % 2 else 1.01))
        N = P.lognormal_ic(x, x[n // 3], 1.0, 1e12, noise=0.05, seed=i)
        b0 = 1e-17 / (x[n // 4] ** 2 if kernel == 2 else 1.0)
        inst.append(dict(n=n, offset=off, beta0=b0, beta_br=6.73e-18,
                         beta_sh=4.0 / 3.0 * float(rng.uniform(10, 500)),
                         log_x0=0.0 if general else float(np.log(x[0])),
                         inv_log_r=0.0 if general else float(1.0 / np.log(x[1] / x[0]))))
        xs.append(x)
        Ns.append(N)
        off += n
    x, N = np.concatenate(xs), np.concatenate(Ns)
    plan = pbe.BatchPlan(inst, kernel_type=kernel, selection=-1, n_quad=0, tile_u=tile_u)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    rhs = hip.DeviceArray(len(N))
    rhs.zero()
    hip.check(plan.launch(dN, dx, None, None, rhs))
    hip.synchronize()
    got = rhs.download()
    plan.destroy()
    for s in inst:
        lo, hi = s["offset"], s["offset"] + s["n"]
        agg = dict(kernel=kernel, beta0=s["beta0"], beta_br=s["beta_br"], beta_sh=s["beta_sh"],
                   log_x0=s["log_x0"], inv_log_r=s["inv_log_r"])
        ra, Aa, _ = oracle_mod.aggregation_rhs(N[lo:hi], x[lo:hi], **agg)
        assert_parity(got[lo:hi], ra, Aa, what=f"tile_u={tile_u} n={s['n']}")
