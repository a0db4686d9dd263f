"""GPU parity of the C-ABI entry points against the CPU oracle.

Every case runs the HIP library through pbe_launch_aggregation_rhs /
pbe_launch_breakage_rhs on seeded inputs and compares with the oracle on the
same arrays: per entry |gpu - oracle| <= 1e-12 * A_i (A_i = sum of absolute
contributions to bin i; bins with A_i = 0 must be exactly 0) and the first
moment sum_i x_i rhs_i within 1e-13 * sum_i x_i A_i of the oracle's exact
clip/daughter defect.
"""
import os

import numpy as np
import pytest

from parity import assert_m1, assert_parity

pytestmark = pytest.mark.gpu

# host threads for the exact oracle at large n (the GPU box's CPU share is 16)
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))


def _gpu(pbe, N, x, agg=None, brk=None, t_q=None, bw_q=None):
    return pbe.rhs_on_device(N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)


def _oracle(orc, N, x, agg=None, brk=None, t_q=None, bw_q=None, threads=1):
    ref = np.zeros(len(N))
    A = np.zeros(len(N))
    D = 0.0
    if agg is not None:
        r, a, d = orc.aggregation_rhs(N, x, threads=threads, **agg)
        ref += r
        A += a
        D += d
    if brk is not None:
        r, a, d = orc.breakage_rhs(N, x, t_q, bw_q, **brk)
        ref += r
        A += a
        D += d
    return ref, A, D


def test_smoluchowski_single_rhs(pbe, oracle_mod):
    from pbe_hip import problems as P
    x, N, agg = P.config_smoluchowski(256)
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    print("n=256 constant", assert_parity(got, ref, A, what="smoluchowski"))
    assert_m1(got, x, A, D, what="smoluchowski")
    # discrete identity of the constant kernel: sum rhs = -beta0 M0^2 / 2
    M0 = N.sum()
    assert abs(got.sum() + 0.5e-12 * M0 * M0) <= 1e-13 * (0.5e-12 * M0 * M0 * 3)


@pytest.mark.parametrize("n", [1024, 4096])
@pytest.mark.parametrize("daughter", ["symmetric", "uniform"])
def test_flocculation_agg_brk(pbe, oracle_mod, n, daughter):
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(n, daughter=daughter, n_quad=16)
    got = _gpu(pbe, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    w = assert_parity(got, ref, A, what=f"n={n} {daughter}")
    m = assert_m1(got, x, A, D, what=f"n={n} {daughter}")
    print(f"n={n} {daughter}: entry {w:.2e} m1 {m:.2e}")


@pytest.mark.parametrize("kernel", range(6))
def test_all_kernels_geometric(pbe, oracle_mod, kernel):
    from pbe_hip import problems as P
    n = 1500
    x = P.geometric_grid(n, 1e-18, 1.02)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=7 + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17 if kernel < 3 else 0.0,
               beta_br=6.73e-18, beta_sh=1.0e-3, log_x0=log_x0, inv_log_r=inv_log_r)
    if kernel == 2:
        agg["beta0"] = 1e-17 / x[n // 3]
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    assert_parity(got, ref, A, what=f"kernel {kernel}")
    assert_m1(got, x, A, D, what=f"kernel {kernel}")


@pytest.mark.parametrize("kernel", range(6))
def test_all_kernels_jittered_binary_search(pbe, oracle_mod, kernel):
    """Non-geometric grid, inv_log_r = 0: the binary-search bracket path."""
    from pbe_hip import problems as P
    n = 3840
    x = P.jittered_grid(n, 1e-18, 1.01, 0.2, seed=3840)
    N = P.lognormal_ic(x, x[n // 4], 2.0, 1e12, noise=0.05, seed=11)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1.0e-3,
               log_x0=0.0, inv_log_r=0.0)
    if kernel == 2:
        agg["beta0"] = 1e-17 / x[n // 4]
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    assert_parity(got, ref, A, what=f"jittered kernel {kernel}")
    assert_m1(got, x, A, D, what=f"jittered kernel {kernel}")


DAUGHTERS = [("symmetric", None), ("erosion", None), ("uniform", 8), ("uniform", 16),
             ("uniform", 32), ("powerlaw", 8), ("powerlaw", 16), ("powerlaw", 32)]


@pytest.mark.parametrize("selection", range(4))
@pytest.mark.parametrize("daughter,nq", DAUGHTERS)
def test_breakage_jittered(pbe, oracle_mod, selection, daughter, nq):
    from pbe_hip import problems as P
    n = 3840
    x = P.jittered_grid(n, 1e-18, 1.01, 0.2, seed=3840)
    N = P.lognormal_ic(x, x[n // 2], 3.0, 1e12, noise=0.05, seed=5)
    t_q, bw_q = P.daughter_table(daughter, nq, q=2.0, eps=0.05)
    brk = dict(selection=selection, S0=1e-3, v_ref=x[n // 2], alpha=1.0 / 3.0,
               v_min=x[n // 2])
    got = _gpu(pbe, N, x, brk=brk, t_q=t_q, bw_q=bw_q)
    ref, A, D = _oracle(oracle_mod, N, x, brk=brk, t_q=t_q, bw_q=bw_q)
    assert_parity(got, ref, A, what=f"sel {selection} {daughter}{nq}")
    assert_m1(got, x, A, D, what=f"sel {selection} {daughter}{nq}")


@pytest.mark.parametrize("nq", [3, 7, 24, 100])
@pytest.mark.parametrize("grid", ["geometric", "jittered"])
def test_breakage_quadrature_sizes(pbe, oracle_mod, nq, grid):
    """Uniform daughter tables whose point count does not divide the
    breakage launch's thread count (a thread's items then change quadrature
    point, and the bracket cursor restarts) or exceeds the 64 points held in
    LDS (read from global memory), linear selection so that every parent
    breaks, on a geometric grid (the cursor's guesses hold) and a jittered
    one (they miss and the search runs); the plain call and the batched form
    (256-thread breakage blocks) against the oracle."""
    from pbe_hip import hip, problems as P
    n = 3000
    x = (P.geometric_grid(n, 1e-18, 1.01) if grid == "geometric"
         else P.jittered_grid(n, 1e-18, 1.01, 0.2, seed=nq))
    N = P.lognormal_ic(x, x[n // 2], 3.0, 1e12, noise=0.05, seed=nq + 1)
    t_q, bw_q = P.daughter_uniform(nq)
    brk = dict(selection=1, S0=1e-3, v_ref=x[n // 2], alpha=1.0, v_min=0.0)
    ref, A, D = _oracle(oracle_mod, N, x, brk=brk, t_q=t_q, bw_q=bw_q)
    got = _gpu(pbe, N, x, brk=brk, t_q=t_q, bw_q=bw_q)
    assert_parity(got, ref, A, what=f"{grid} n_quad {nq}")
    assert_m1(got, x, A, D, what=f"{grid} n_quad {nq}")
    inst = [dict(n=n, offset=0, S0=brk["S0"], v_ref=brk["v_ref"], alpha=brk["alpha"],
                 v_min=brk["v_min"])]
    plan = pbe.BatchPlan(inst, kernel_type=-1, selection=1, n_quad=nq)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
    drhs = hip.DeviceArray(n)
    drhs.zero()
    hip.check(plan.launch(dN, dx, dt, db, drhs), "pbe_launch_batched_rhs")
    hip.synchronize()
    assert_parity(drhs.download(), ref, A, what=f"batched {grid} n_quad {nq}")


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 127, 129, 255, 257, 511, 513, 1000])
def test_ragged_sizes(pbe, oracle_mod, n):
    """Sizes around the tile edges, including n = 1 and 2 where every pair clips."""
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
    got = _gpu(pbe, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    assert_parity(got, ref, A, what=f"n={n}")


def test_negative_and_zero_populations(pbe, oracle_mod):
    """Aggregation has no guard (negative N contributes); breakage skips
    N <= 2.2e-308 (src/breakage_kernels.cuh:131,195)."""
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
    got = _gpu(pbe, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    assert_parity(got, ref, A, what="negative N")


def test_accumulates_into_rhs(pbe, oracle_mod):
    """rhs is accumulated into, not overwritten (aggregation.cuh:88-91)."""
    from pbe_hip import hip, problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(1024)
    n = len(N)
    base = np.linspace(-1.0, 1.0, n) * 1e3
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray.from_host(base)
    dt, db = hip.DeviceArray.from_host(t_q), hip.DeviceArray.from_host(bw_q)
    hip.check(pbe.launch_breakage_rhs(dN, dx, dt, db, drhs,
                                      pbe.brk_params(n, len(t_q), brk)))
    hip.check(pbe.launch_aggregation_rhs(dN, dx, drhs, pbe.agg_params(n, agg)))
    hip.synchronize()
    got = drhs.download()
    ref, A, _ = _oracle(oracle_mod, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    # each atomic add into a bin holding ~|base| rounds at ulp(|base|)
    slack = 64 * np.spacing(np.abs(base) + np.abs(ref))
    assert np.all(np.abs(got - base - ref) <= 1e-12 * A + slack)


@pytest.mark.parametrize("daughter", ["symmetric", "uniform"])
def test_flocculation_16384(pbe, oracle_mod, daughter):
    """Full-size workload (n = 16384, r = 1.02, Brownian+shear, threshold
    breakage) with symmetric-binary daughters and with the 16-point
    Gauss-Legendre uniform table: every entry against the exact oracle at
    1e-12 A_i and the first-moment identity at 1e-13, and against the OpenMP
    port (plain FP64, the reference's bracket rule)."""
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(16384, daughter=daughter, n_quad=16)
    got = _gpu(pbe, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q)
    port = oracle_mod.port_aggregation_rhs(N, x, threads=ORACLE_THREADS, **agg)
    port += oracle_mod.port_breakage_rhs(N, x, t_q, bw_q, threads=ORACLE_THREADS, **brk)
    scale = np.abs(port).max()
    assert np.all(np.abs(got - port) <= 1e-10 * scale + 1e-9 * np.abs(port))
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, brk=brk, t_q=t_q, bw_q=bw_q,
                        threads=ORACLE_THREADS)
    w = assert_parity(got, ref, A, what=f"n=16384 {daughter}")
    m = assert_m1(got, x, A, D, what=f"n=16384 {daughter}")
    print(f"n=16384 {daughter} (n_quad {len(t_q)}): entry {w:.2e} m1 {m:.2e}")


@pytest.mark.parametrize("n,r,search", [(32768, 1.01, False), (40007, 1.008, True),
                                        (65536, 1.005, False)])
def test_beyond_int32_pair_count(pbe, oracle_mod, n, r, search):
    """Sizes past the reference's n = 23,169 (its triangle decode computes
    (2n+1)^2 in int, src/aggregation_kernels.cuh:171): 5.4e8 to 2.1e9 pairs
    through the band sweep and the column strips (64-bit plan).  Geometric
    grids with the log-ratio lookup, and a ragged n (157 column blocks, the
    last partial) with the binary search.  Aggregation and breakage are
    checked apart (their magnitudes differ by 35 decades here), each entry
    against the exact oracle at 1e-12 A_i and the first moment at 1e-13.
    n = 65536 at r = 1.005 has a near-pair reach of 1062 bins, so the strips
    hold per-pair near chunks with brackets beyond their LDS windows."""
    from pbe_hip import problems as P
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 4], 0.8, 1e12, noise=0.05, seed=n)
    log_x0, inv_log_r = (0.0, 0.0) if search else P.grid_lookup(x)
    agg = dict(kernel=5, beta0=0.0, beta_br=6.73e-18, beta_sh=4.0 / 3.0 * 100.0,
               log_x0=log_x0, inv_log_r=inv_log_r)
    brk = dict(selection=3, S0=1e-3, v_ref=1.0, alpha=1.0, v_min=x[n // 4])
    t_q, bw_q = P.daughter_symmetric()
    ga = _gpu(pbe, N, x, agg=agg)
    gb = _gpu(pbe, N, x, brk=brk, t_q=t_q, bw_q=bw_q)
    ra, Aa, Da = _oracle(oracle_mod, N, x, agg=agg, threads=ORACLE_THREADS)
    rb, Ab, Db = _oracle(oracle_mod, N, x, brk=brk, t_q=t_q, bw_q=bw_q)
    wa = assert_parity(ga, ra, Aa, what=f"aggregation n={n}")
    ma = assert_m1(ga, x, Aa, Da, what=f"aggregation n={n}")
    wb = assert_parity(gb, rb, Ab, what=f"breakage n={n}")
    mb = assert_m1(gb, x, Ab, Db, what=f"breakage n={n}")
    print(f"n={n}: aggregation entry {wa:.2e} m1 {ma:.2e}; breakage entry {wb:.2e} m1 {mb:.2e}")


@pytest.mark.parametrize("n,r,kernel", [(9000, 1.002, 5), (9000, 1.005, 5), (9000, 1.005, 0),
                                         (20000, 1.004, 3)])
def test_wide_strips_fine_grid(pbe, oracle_mod, n, r, kernel):
    """512-column blocks (tile_u = 8) on geometric grids whose near-pair reach
    (1062 bins at r = 1.005, 2656 at 1.002) exceeds the block and its LDS
    window: strip chunks take the per-pair near path with brackets found
    beyond the 768-bin window (log-ratio guess plus fix-up).  Every entry
    against the exact oracle."""
    from pbe_hip import hip, problems as P
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=n + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    drhs.zero()
    hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, pbe.agg_params(n, agg), tile_u=8,
                                            band_split=0))
    hip.synchronize()
    got = drhs.download()
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, threads=ORACLE_THREADS)
    what = f"U=8 n={n} r={r} kernel {kernel}"
    assert_parity(got, ref, A, what=what)
    assert_m1(got, x, A, D, what=what)


@pytest.mark.parametrize("u", [1, 2, 4, 8])
@pytest.mark.parametrize("n,kernel", [(700, 5), (3000, 5), (3000, 0), (1200, 2)])
def test_every_tile_shape(pbe, oracle_mod, u, n, kernel):
    """Each compiled tile shape (64*U square tiles) through
    pbe_launch_aggregation_rhs_ex gives the same parity."""
    from pbe_hip import hip, problems as P
    x = P.geometric_grid(n, 1e-18, 1.02)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=u)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 3],
               beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0, inv_log_r=inv_log_r)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    drhs.zero()
    p = pbe.agg_params(n, agg)
    hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u, band_split=16 // u))
    hip.synchronize()
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    got = drhs.download()
    assert_parity(got, ref, A, what=f"U={u} n={n}")
    assert_m1(got, x, A, D, what=f"U={u} n={n}")


@pytest.mark.parametrize("n,r,kernel", [(4096, 1.02, 5), (6000, 1.01, 0), (3500, 1.03, 3),
                                         (5000, 1.005, 5)])
def test_tile_launch_form(pbe, oracle_mod, n, r, kernel):
    """Launch form 3 (one launch of every tile, the tile diagonals within the
    near-pair reach as row slices) at mid sizes, where the tuned table picks
    it, at every tile width and two band splits."""
    from pbe_hip import hip, problems as P
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 4], 0.8, 1e12, noise=0.05, seed=n + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 4],
               beta_br=6.73e-18, beta_sh=4.0 / 3.0 * 100.0, log_x0=log_x0, inv_log_r=inv_log_r)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, threads=ORACLE_THREADS)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    p = pbe.agg_params(n, agg)
    for u in (1, 2, 4, 8):
        for split in (2, 8):
            drhs.zero()
            hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u, band_split=split,
                                                    form=3))
            hip.synchronize()
            got = drhs.download()
            what = f"form 3 U={u} split {split} n={n} r={r} kernel {kernel}"
            assert_parity(got, ref, A, what=what)
            assert_m1(got, x, A, D, what=what)


@pytest.mark.parametrize("r", [1.005, 1.002])
@pytest.mark.parametrize("u,n", [(1, 384), (1, 1500), (2, 1500), (4, 900)])
@pytest.mark.parametrize("kernel", [0, 5])
def test_tiles_sliced_near_diagonals(pbe, oracle_mod, r, u, n, kernel):
    """The single tile launch on grids whose near-pair reach spans many tiles
    (r = 1.005: 1062 bins, r = 1.002: 3107): every tile diagonal within the
    reach runs as row slices, the rest whole, column n-1 through the far
    loop."""
    from pbe_hip import hip, problems as P
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 2], 1.5, 1e12, noise=0.05, seed=u + n)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    drhs.zero()
    p = pbe.agg_params(n, agg)
    hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u, band_split=8, form=1))
    hip.synchronize()
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    got = drhs.download()
    assert_parity(got, ref, A, what=f"U={u} n={n} r={r}")
    assert_m1(got, x, A, D, what=f"U={u} n={n} r={r}")


def test_strips_jittered_grid(pbe, oracle_mod):
    """Large non-geometric grid (r = 1.01 +-20 %): the column-strip launch with
    the binary-search bracket and near pairs below the diagonal band."""
    from pbe_hip import problems as P
    n = 12000
    x = P.jittered_grid(n, 1e-18, 1.01, 0.2, seed=12000)
    N = P.lognormal_ic(x, x[n // 3], 3.0, 1e12, noise=0.05, seed=13)
    agg = dict(kernel=5, beta0=0.0, beta_br=6.73e-18, beta_sh=1.0e-3, log_x0=0.0,
               inv_log_r=0.0)
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    assert_parity(got, ref, A, what="strips jittered n=12000")
    assert_m1(got, x, A, D, what="strips jittered n=12000")


@pytest.mark.parametrize("jitter,kernel", [(0.05, 5), (0.3, 5), (0.3, 0)])
def test_geometric_lookup_on_jittered_grid(pbe, oracle_mod, jitter, kernel):
    """A grid declared geometric (log_x0, inv_log_r of its nominal ratio 1.02)
    whose section boundaries are jittered by up to `jitter` of a step: the
    log-ratio guesses and the band's start-bracket offsets delta(e) then miss
    in many lanes, and each miss must be caught against the pivots and fixed
    up, so every pair still lands in its exact bracket (the oracle's exact
    bracket mode)."""
    from pbe_hip import problems as P
    n = 9000
    x = P.jittered_grid(n, 1e-18, 1.02, jitter, seed=int(jitter * 100) + kernel)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=77 + kernel)
    log_x0, inv_log_r = np.log(1e-18 * np.sqrt(1.02)), 1.0 / np.log(1.02)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    assert_parity(got, ref, A, what=f"declared geometric, jitter {jitter}, kernel {kernel}")
    assert_m1(got, x, A, D, what=f"declared geometric, jitter {jitter}, kernel {kernel}")


@pytest.mark.parametrize("r,kernel", [(1.001, 5), (1.001, 0), (1.005, 3)])
def test_strips_fine_grid(pbe, oracle_mod, r, kernel):
    """Geometric grids finer than the band (near reach 1000+ bins at
    r = 1.001): every strip chunk takes the per-pair near path and most
    brackets lie beyond the LDS windows (log-ratio guess plus fix-up)."""
    from pbe_hip import problems as P
    n = 5000
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 2], 0.5, 1e12, noise=0.05, seed=int(r * 1e4) + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    assert_parity(got, ref, A, what=f"fine grid r={r} kernel {kernel}")
    assert_m1(got, x, A, D, what=f"fine grid r={r} kernel {kernel}")


@pytest.mark.parametrize("kernel", range(6))
def test_fused_sweep_and_strips(pbe, oracle_mod, kernel):
    """Launch forms 2 and 4 (at most 64 blocks of 256 columns; two or three
    workgroups per CU): the band sweep items and the four-wave strip shares
    in one launch, for every band split (the launch halves it while the
    sweep would take over half the slots), on geometric grids (start-bracket
    offsets; r = 1.01 puts near chunks in the strips) and a jittered one
    (searches); n = 16385 (65 blocks) takes two launches."""
    from pbe_hip import hip, problems as P
    for n, jitter, r in ((3400, 0.0, 1.02), (5000, 0.0, 1.02), (8192, 0.0, 1.02),
                         (6000, 0.2, 1.02), (12000, 0.0, 1.01), (16384, 0.0, 1.02),
                         (16385, 0.0, 1.02)):
        x = (P.geometric_grid(n, 1e-18, r) if jitter == 0.0
             else P.jittered_grid(n, 1e-18, r, jitter, seed=n))
        N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=60 + kernel)
        log_x0, inv_log_r = P.grid_lookup(x) if jitter == 0.0 else (0.0, 0.0)
        agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 3],
                   beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0, inv_log_r=inv_log_r)
        ref, A, D = _oracle(oracle_mod, N, x, agg=agg, threads=ORACLE_THREADS)
        dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
        drhs = hip.DeviceArray(n)
        p = pbe.agg_params(n, agg)
        for split, form in ((1, 2), (4, 2), (16, 2), (2, 4), (16, 4)):
            drhs.zero()
            hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=4, band_split=split,
                                                    form=form))
            hip.synchronize()
            got = drhs.download()
            what = f"fused form {form} kernel {kernel} n={n} r={r} jitter {jitter} split {split}"
            assert_parity(got, ref, A, what=what)
            assert_m1(got, x, A, D, what=what)


@pytest.mark.parametrize("n,r,kernel", [(6000, 1.015, 5), (8192, 1.01, 0), (8192, 1.01, 5),
                                         (12000, 1.015, 3), (16384, 1.005, 5)])
def test_fine_grid_default_width(pbe, oracle_mod, n, r, kernel):
    """The default launch (the tuned table's shape) on geometric grids whose
    far-pair reach (282 bins at r = 1.015, 463 at 1.01) exceeds 256 columns,
    where the strips hold near chunks."""
    from pbe_hip import problems as P
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=90 + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17, beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0,
               inv_log_r=inv_log_r)
    got = _gpu(pbe, N, x, agg=agg)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg, threads=ORACLE_THREADS)
    assert_parity(got, ref, A, what=f"default width n={n} r={r} kernel {kernel}")
    assert_m1(got, x, A, D, what=f"default width n={n} r={r} kernel {kernel}")


@pytest.mark.parametrize("kernel", range(6))
def test_band_every_split(pbe, oracle_mod, kernel):
    """n = 9000: the column strips plus the diagonal band, for every band
    split and the 128-, 256- and 512-column blocks."""
    from pbe_hip import hip, problems as P
    n = 9000
    x = P.geometric_grid(n, 1e-18, 1.02)
    N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=40 + kernel)
    log_x0, inv_log_r = P.grid_lookup(x)
    agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 3],
               beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0, inv_log_r=inv_log_r)
    ref, A, D = _oracle(oracle_mod, N, x, agg=agg)
    dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
    drhs = hip.DeviceArray(n)
    p = pbe.agg_params(n, agg)
    for u, form in ((2, 1), (4, 1), (4, 2), (8, 1)):
        for split in (1, 2, 4, 8, 16):
            drhs.zero()
            hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u, band_split=split,
                                                    form=form))
            hip.synchronize()
            got = drhs.download()
            what = f"kernel {kernel} U={u} split {split} form {form}"
            assert_parity(got, ref, A, what=what)
            assert_m1(got, x, A, D, what=what)


@pytest.mark.parametrize("kernel", [0, 3, 5])
def test_zero_birth_rows(pbe, oracle_mod, kernel):
    """Coarse geometric grids (r = 1.05, 1.1) where most strip rows lie below
    half an ulp of their block's first pivot: x_j + x_k rounds to x_k, the
    far-pair birth weight is exactly 0 and the strips only add those rows'
    deaths.  Every strip launch shape, a jittered grid (declared not
    geometric: the cost line prices no zero-birth chunks, the kernel still
    skips them) and populations with random empty bins (bins fed only by
    such pairs must stay exact)."""
    from pbe_hip import hip, problems as P
    rng = np.random.default_rng(70 + kernel)
    for n, r, jitter, holes in ((6000, 1.05, 0.0, False), (3000, 1.1, 0.0, True),
                                (6000, 1.05, 0.2, True)):
        x = (P.geometric_grid(n, 1e-18, r) if jitter == 0.0
             else P.jittered_grid(n, 1e-18, r, jitter, seed=n))
        N = P.lognormal_ic(x, x[n // 3], 1.5, 1e12, noise=0.05, seed=80 + kernel)
        # a broad floor keeps the far rows populated (the lognormal alone
        # underflows there)
        N = N + 1e-3 * rng.random(n)
        if holes:
            N[rng.random(n) < 0.3] = 0.0
        log_x0, inv_log_r = P.grid_lookup(x) if jitter == 0.0 else (0.0, 0.0)
        agg = dict(kernel=kernel, beta0=1e-17 if kernel != 2 else 1e-17 / x[n // 3],
                   beta_br=6.73e-18, beta_sh=1e-3, log_x0=log_x0, inv_log_r=inv_log_r)
        ref, A, D = _oracle(oracle_mod, N, x, agg=agg, threads=ORACLE_THREADS)
        dN, dx = hip.DeviceArray.from_host(N), hip.DeviceArray.from_host(x)
        drhs = hip.DeviceArray(n)
        p = pbe.agg_params(n, agg)
        for u, split, form in ((4, 4, 1), (4, 4, 2), (4, 8, 4), (2, 4, 1), (8, 4, 1), (0, 0, 0)):
            drhs.zero()
            hip.check(pbe.launch_aggregation_rhs_ex(dN, dx, drhs, p, tile_u=u, band_split=split,
                                                    form=form))
            hip.synchronize()
            got = drhs.download()
            what = f"zero-birth rows kernel {kernel} n={n} r={r} jitter {jitter} U={u} form {form}"
            assert_parity(got, ref, A, what=what)
            assert_m1(got, x, A, D, what=what)
