"""GPU tests of the atomic-free gather mode
(pbe_launch_aggregation_rhs_deterministic / pbe_launch_breakage_rhs_deterministic).

Parity against the CPU oracle at the same bar as the atomic launches (per
entry |gpu - oracle| <= 1e-12 * A_i, exact zeros where A_i = 0, first moment
within 1e-13 * sum x_i A_i of the oracle's defect), bitwise identity of
repeated calls, and agreement with the atomic launches to rounding.
"""
import os

import numpy as np
import pytest

from parity import assert_m1, assert_parity

pytestmark = pytest.mark.gpu

# host threads for the exact oracle at large n (the GPU box's CPU share is 16)
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))

REPEATS = 3


def _det(pbe, N, x, agg=None, brk=None, t_q=None, bw_q=None, base=None):
    """One deterministic evaluation (aggregation then breakage) on a fresh
    workspace; returns the rhs."""
    from pbe_hip import hip
    n = len(N)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray.from_host(np.zeros(n) if base is None else base)
    ws_bytes = max(pbe.aggregation_deterministic_workspace_bytes(n),
                   pbe.breakage_deterministic_workspace_bytes(n))
    ws = hip.DeviceArray(ws_bytes // 8 + 1)
    if agg is not None:
        hip.check(pbe.launch_aggregation_rhs_deterministic(dN, dx, drhs, pbe.agg_params(n, agg),
                                                           ws))
    if brk is not None:
        dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
        hip.check(pbe.launch_breakage_rhs_deterministic(
            dN, dx, dt, db, drhs, pbe.brk_params(n, len(t_q), brk), ws))
    hip.synchronize()
    return drhs.download()


def _oracle(orc, N, x, agg=None, brk=None, t_q=None, bw_q=None, threads=1):
    ref, A, D = np.zeros(len(N)), np.zeros(len(N)), 0.0
    if agg is not None:
        r, a, d = orc.aggregation_rhs(N, x, threads=threads, **agg)
        ref, A, D = ref + r, A + a, D + d
    if brk is not None:
        r, a, d = orc.breakage_rhs(N, x, t_q, bw_q, **brk)
        ref, A, D = ref + r, A + a, D + d
    return ref, A, D


def _check(pbe, orc, N, x, what, agg=None, brk=None, t_q=None, bw_q=None, m1=True):
    runs = [_det(pbe, N, x, agg, brk, t_q, bw_q) for _ in range(REPEATS)]
    for r in runs[1:]:
        assert np.array_equal(r.view(np.uint64), runs[0].view(np.uint64)), \
            f"{what}: repeated calls differ"
    got = runs[0]
    ref, A, D = _oracle(orc, N, x, agg, brk, t_q, bw_q)
    w = assert_parity(got, ref, A, what=what)
    if m1:
        assert_m1(got, x, A, D, what=what)
    return got, w


@pytest.mark.parametrize("kernel", range(6))
def test_det_all_kernels_geometric(pbe, oracle_mod, kernel):
    from pbe_hip import problems as P
    n = 1500
    x = P.geometric_grid(n, 1e-18, 1.02)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=7 + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 3],
               beta_br=6.73e-18, beta_sh=1.0e-3, log_x0=log_x0, inv_log_r=inv_log_r)
    got, _ = _check(pbe, oracle_mod, N, x, f"det kernel {kernel}", agg=agg)
    atomic = pbe.rhs_on_device(N, x, agg=agg)
    _, A, _ = _oracle(oracle_mod, N, x, agg=agg)
    assert np.all(np.abs(got - atomic) <= 2e-12 * A + 2.0 ** -1000 * A.max())


@pytest.mark.parametrize("kernel", [0, 3, 5])
def test_det_jittered_grid(pbe, oracle_mod, kernel):
    """Non-geometric grid (r = 1.01 with +-20 % jitter): many near pairs."""
    from pbe_hip import problems as P
    n = 3840
    x = P.jittered_grid(n, 1e-18, 1.01, 0.2, seed=3840)
    N = P.lognormal_ic(x, x[n // 4], 2.0, 1e12, noise=0.05, seed=11)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1.0e-3,
               log_x0=0.0, inv_log_r=0.0)
    _check(pbe, oracle_mod, N, x, f"det jittered kernel {kernel}", agg=agg)


@pytest.mark.parametrize("r", [1.001, 1.3, 2.5])
def test_det_grid_ratios(pbe, oracle_mod, r):
    """Fine (mostly near pairs), coarse and very coarse (all pairs far or
    clipped) geometric grids."""
    from pbe_hip import problems as P
    n = 900 if r < 1.1 else 120
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 3], 1.0, 1e12, noise=0.05, seed=3)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=5, beta0=0.0, beta_br=6.73e-18, beta_sh=1.0e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    t_q, bw_q = P.daughter_uniform(8)
    brk = dict(selection=0, S0=1e-3, v_ref=1.0, alpha=1.0, v_min=0.0)
    _check(pbe, oracle_mod, N, x, f"det r={r}", agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)


DAUGHTERS = [("symmetric", None), ("erosion", None), ("uniform", 16), ("powerlaw", 16)]


@pytest.mark.parametrize("selection", range(4))
@pytest.mark.parametrize("daughter,nq", DAUGHTERS)
def test_det_breakage_jittered(pbe, oracle_mod, selection, daughter, nq):
    from pbe_hip import problems as P
    n = 3840
    x = P.jittered_grid(n, 1e-18, 1.01, 0.2, seed=3840)
    N = P.lognormal_ic(x, x[n // 2], 3.0, 1e12, noise=0.05, seed=5)
    t_q, bw_q = P.daughter_table(daughter, nq, q=2.0, eps=0.05)
    brk = dict(selection=selection, S0=1e-3, v_ref=x[n // 2], alpha=1.0 / 3.0,
               v_min=x[n // 2])
    _check(pbe, oracle_mod, N, x, f"det sel {selection} {daughter}{nq}", brk=brk, t_q=t_q,
           bw_q=bw_q)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 129, 257, 1000, 1025])
def test_det_ragged_sizes(pbe, oracle_mod, n):
    from pbe_hip import problems as P
    rng = np.random.default_rng(n)
    x = P.geometric_grid(n, 1e-18, 1.05)
    N = rng.uniform(0.0, 1e12, n)
    log_x0 = np.log(x[0])
    inv_log_r = 1.0 / np.log(1.05) if n > 1 else 0.0
    agg = dict(kernel=5, beta0=0.0, beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    t_q, bw_q = P.daughter_uniform(8)
    brk = dict(selection=2, S0=1e-3, v_ref=x[n // 2], alpha=0.5, v_min=0.0)
    _check(pbe, oracle_mod, N, x, f"det n={n}", agg=agg, brk=brk, t_q=t_q, bw_q=bw_q,
           m1=False)


def test_det_negative_and_zero_populations(pbe, oracle_mod):
    from pbe_hip import problems as P
    n = 700
    rng = np.random.default_rng(1)
    x = P.geometric_grid(n, 1e-18, 1.03)
    N = rng.normal(0.0, 1e10, n)
    N[rng.uniform(size=n) < 0.3] = 0.0
    N[5] = 1e-310
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=4, beta0=0.0, beta_br=0.0, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    t_q, bw_q = P.daughter_erosion(0.05)
    brk = dict(selection=1, S0=1e-3, v_ref=x[n // 2], alpha=1.0, v_min=0.0)
    _check(pbe, oracle_mod, N, x, "det negative N", agg=agg, brk=brk, t_q=t_q, bw_q=bw_q,
           m1=False)


def test_det_accumulates(pbe, oracle_mod):
    """rhs += source: one add per bin onto the caller's values."""
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(1024)
    base = np.linspace(-1.0, 1.0, len(N)) * 1e3
    got = _det(pbe, N, x, agg, brk, t_q, bw_q, base=base)
    ref, A, _ = _oracle(oracle_mod, N, x, agg, brk, t_q, bw_q)
    slack = 4 * np.spacing(np.abs(base) + np.abs(ref))
    assert np.all(np.abs(got - base - ref) <= 1e-12 * A + slack)


def test_det_flocculation_16384(pbe, oracle_mod):
    """The bench workload: parity with the oracle and bitwise repeatability."""
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(16384)
    _, w = _check(pbe, oracle_mod, N, x, "det n=16384", agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    print(f"det n=16384: entry {w:.2e}")


def test_det_beyond_int32_pair_count(pbe, oracle_mod):
    """n = 65536 (2.1e9 pairs, past the reference's n = 23,169 int limit):
    bitwise repeatability; every entry against the exact oracle (1e-12 A_i)
    with the first moment, aggregation and breakage apart; and every entry
    against the atomic launches (their sums differ in order only)."""
    from pbe_hip import problems as P
    n = 65536
    x = P.geometric_grid(n, 1e-18, 1.005)
    N = P.lognormal_ic(x, x[n // 4], 0.8, 1e12, noise=0.05, seed=n)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=5, beta0=0.0, beta_br=6.73e-18, beta_sh=4.0 / 3.0 * 100.0,
               log_x0=log_x0, inv_log_r=inv_log_r)
    brk = dict(selection=3, S0=1e-3, v_ref=1.0, alpha=1.0, v_min=x[n // 4])
    t_q, bw_q = P.daughter_symmetric()
    for what, kw in (("aggregation", dict(agg=agg)),
                     ("breakage", dict(brk=brk, t_q=t_q, bw_q=bw_q))):
        runs = [_det(pbe, N, x, **kw) for _ in range(2)]
        assert np.array_equal(runs[0].view(np.uint64), runs[1].view(np.uint64)), \
            f"det {what} n={n}: repeated calls differ"
        got = runs[0]
        ref, A, D = _oracle(oracle_mod, N, x, threads=ORACLE_THREADS, **kw)
        w = assert_parity(got, ref, A, what=f"det {what} n={n}")
        m = assert_m1(got, x, A, D, what=f"det {what} n={n}")
        atomic = pbe.rhs_on_device(N, x, **kw)
        assert_parity(got, atomic, A, tol=2e-12, what=f"det vs atomic {what} n={n}")
        print(f"det {what} n={n}: entry {w:.2e} m1 {m:.2e}")


def test_det_workspace_checks(pbe):
    """A missing or short workspace is refused before any launch."""
    from pbe_hip import hip, problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(256)
    n = len(N)
    dN, dx, drhs = (hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x),
                    hip.DeviceArray(n))
    small = hip.DeviceArray(4)
    assert pbe.launch_aggregation_rhs_deterministic(dN, dx, drhs, pbe.agg_params(n, agg),
                                                    small) == 1
    assert pbe.launch_aggregation_rhs_deterministic(dN, dx, drhs, pbe.agg_params(n, agg),
                                                    None) == 17
