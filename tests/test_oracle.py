"""CPU tests that pin the oracle before it is trusted as the checker.

Pins, each chosen so that a plausible slip in the oracle (dropped term, wrong
sign, wrong bracket index, swapped split weights) fails at least one:
  * golden vectors from an independent 50-digit mpmath restatement written
    with hat functions and a full-row death sum (tests/golden);
  * exact discrete identities of the semi-discrete fixed-pivot system;
  * the reference's own known answers and invariants
    (tests/unit/test_aggregation.cpp, tests/unit/test_breakage.cpp) and the
    errors its five example programs document (docs/examples.md);
  * the Smoluchowski closed form for the constant kernel under RK4, and the
    sum-kernel, product-kernel and linear-selection moment laws under RK4;
  * the reference bracket rule (log-ratio formula) against the exact bracket.
"""
import json
import math
import os

import numpy as np
import pytest

from parity import assert_parity

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "fixed_pivot_golden.json")


def _golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


CASES = _golden_cases()


@pytest.mark.parametrize("case", CASES, ids=[
    f"{c['kind'][:3]}-{c['grid']}-" + (f"k{c['kernel']}" if c["kind"] == "aggregation"
                                       else f"{c['daughter']}-s{c['selection']}")
    for c in CASES])
def test_oracle_matches_golden(oracle_mod, case):
    x, N = np.array(case["x"]), np.array(case["N"])
    if case["kind"] == "aggregation":
        for mode in (oracle_mod.BRACKET_EXACT, oracle_mod.BRACKET_REFERENCE):
            rhs, A, _ = oracle_mod.aggregation_rhs(
                N, x, case["kernel"], case["beta0"], case["beta_br"], case["beta_sh"],
                case["log_x0"], case["inv_log_r"], bracket=mode)
            assert_parity(rhs, case["rhs"], case["A"], tol=1e-14, what="golden agg")
            np.testing.assert_allclose(A, case["A"], rtol=1e-13, atol=0)
    else:
        rhs, A, _ = oracle_mod.breakage_rhs(
            N, x, case["t_q"], case["bw_q"], case["selection"], case["S0"], case["v_ref"],
            case["alpha"], case["v_min"])
        assert_parity(rhs, case["rhs"], case["A"], tol=1e-14, what="golden brk")
        np.testing.assert_allclose(A, case["A"], rtol=1e-13, atol=0)


# ------------------------------------------------------------- identities --

def _grid(n=300, r=1.04):
    from pbe_hip import problems as P
    x = P.geometric_grid(n, 1e-18, r)
    N = P.lognormal_ic(x, x[n // 3], 1.2, 1e12, noise=0.05, seed=n)
    return x, N, *P.grid_lookup(x)


def test_constant_kernel_number_identity(oracle_mod):
    x, N, lx, ilr = _grid()
    rhs, A, _ = oracle_mod.aggregation_rhs(N, x, 0, 1e-12, 0, 0, lx, ilr)
    M0 = N.sum()
    assert abs(rhs.sum() + 0.5e-12 * M0 * M0) <= 1e-13 * 0.5e-12 * M0 * M0


def test_sum_kernel_number_identity(oracle_mod):
    x, N, lx, ilr = _grid()
    b0 = 1e-12 / x[100]
    rhs, A, _ = oracle_mod.aggregation_rhs(N, x, 1, b0, 0, 0, lx, ilr)
    M0, M1 = N.sum(), (x * N).sum()
    assert abs(rhs.sum() + b0 * M0 * M1) <= 1e-12 * b0 * M0 * M1


def test_product_kernel_number_identity(oracle_mod):
    x, N, lx, ilr = _grid()
    b0 = 1e-12 / x[100] ** 2
    rhs, A, _ = oracle_mod.aggregation_rhs(N, x, 2, b0, 0, 0, lx, ilr)
    M1 = (x * N).sum()
    assert abs(rhs.sum() + 0.5 * b0 * M1 * M1) <= 1e-12 * 0.5 * b0 * M1 * M1


@pytest.mark.parametrize("kernel", range(6))
def test_aggregation_volume_identity(oracle_mod, kernel):
    """sum_i x_i rhs_i equals the exact clip defect (zero without clipping)."""
    x, N, lx, ilr = _grid(200, 1.08)
    rhs, A, D = oracle_mod.aggregation_rhs(N, x, kernel, 1e-17 / x[60] ** (kernel == 2),
                                           6.73e-18, 1e-3, lx, ilr)
    assert abs((x * rhs).sum() - D) <= 1e-13 * (x * A).sum()
    assert D < 0.0  # the grid is narrow enough that large pairs clip at n-1


@pytest.mark.parametrize("daughter,nq", [("symmetric", None), ("erosion", None),
                                         ("uniform", 12), ("powerlaw", 12)])
@pytest.mark.parametrize("sel", range(4))
def test_breakage_identities(oracle_mod, daughter, nq, sel):
    from pbe_hip import problems as P
    x, N, lx, ilr = _grid(250, 1.05)
    t_q, bw_q = P.daughter_table(daughter, nq, q=2.0, eps=0.05)
    kw = dict(S0=1e-3, v_ref=x[125], alpha=0.4, v_min=x[125])
    rhs, A, D = oracle_mod.breakage_rhs(N, x, t_q, bw_q, sel, **kw)
    S = np.array([oracle_mod.selection(sel, v, **kw) for v in x])
    SN = np.where((S != 0) & (N > 2.2e-308), S * N, 0.0)
    # number: sum rhs = sum_j S_j N_j (sum_q bw_q - 1), clipping conserves number
    expect = SN.sum() * (bw_q.sum() - 1.0)
    assert abs(rhs.sum() - expect) <= 1e-12 * (A.sum())
    # volume: sum x rhs = defect (daughter moment + clipping)
    assert abs((x * rhs).sum() - D) <= 1e-13 * (x * A).sum()


def test_bracket_rules_agree_on_flocculation_grid(oracle_mod):
    """The reference's log-ratio bracket and the exact bracket coincide on the
    r = 1.02 grid (log rounding never crosses an integer there)."""
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(1024)
    a = oracle_mod.aggregation_rhs(N, x, bracket=oracle_mod.BRACKET_EXACT, **agg)
    b = oracle_mod.aggregation_rhs(N, x, bracket=oracle_mod.BRACKET_REFERENCE, **agg)
    assert np.array_equal(a[0], b[0])


# ---------------------------------------- reference tests, on the oracle --

R64 = 1.122018


def _ref_fixture():
    x = [1.0e-18]
    for _ in range(63):
        x.append(x[-1] * R64)
    x = np.array(x)
    return x, math.log(x[0]), 1.0 / math.log(x[1] / x[0])


def test_reference_self_collision_known_answer(oracle_mod):
    """test_aggregation.cpp:251-266: rhs[32] = -beta0 Nj^2 = -1e11."""
    x, lx, ilr = _ref_fixture()
    N = np.zeros(64)
    N[32] = 1e14
    rhs, _, _ = oracle_mod.aggregation_rhs(N, x, 0, 1e-17, 0, 0, lx, ilr)
    assert abs(rhs[32] - (-1e11)) <= 1e11 * 1e-10
    assert abs(rhs.sum() + 0.5e11) <= 1e-3


def test_reference_death_constant_selection(oracle_mod):
    """test_breakage.cpp:241-254: rhs[32] = -S0 Nj = -1e11."""
    x, _, _ = _ref_fixture()
    N = np.zeros(64)
    N[32] = 1e14
    rhs, _, _ = oracle_mod.breakage_rhs(N, x, [0.5], [2.0], 0, 1e-3)
    assert abs(rhs[32] - (-1e11)) <= 1e11 * 1e-10


def test_reference_zero_population(oracle_mod):
    x, lx, ilr = _ref_fixture()
    N = np.zeros(64)
    for k in range(6):
        rhs, _, _ = oracle_mod.aggregation_rhs(N, x, k, 1e-17, 1e-17, 1e-17, lx, ilr)
        assert np.all(rhs == 0.0)
    rhs, _, _ = oracle_mod.breakage_rhs(N, x, [0.5], [2.0], 0, 1e-3)
    assert np.all(rhs == 0.0)


def test_reference_sign_invariants(oracle_mod):
    x, lx, ilr = _ref_fixture()
    N = np.full(64, 1e12)
    rhs, _, _ = oracle_mod.aggregation_rhs(N, x, 0, 1e-17, 0, 0, lx, ilr)
    assert rhs.sum() <= 0.0
    rhs, _, _ = oracle_mod.breakage_rhs(N, x, [0.5], [2.0], 0, 1e-3)
    assert rhs.sum() >= 0.0
    rhs, _, _ = oracle_mod.breakage_rhs(N, x, [0.5], [2.0], 3, 1e-3, 1.0, 1.0, x[32])
    assert np.all(rhs[:32] >= 0.0)


@pytest.mark.parametrize("kernel,b0,bbr,bsh", [(0, 1e-17, 0, 0), (1, 1e-17, 0, 0),
                                               (2, 1e-17, 0, 0), (3, 0, 1e-17, 0),
                                               (4, 0, 0, 1e-17), (5, 0, 1e-17, 1e-17)])
def test_reference_aggregation_conservation(oracle_mod, kernel, b0, bbr, bsh):
    """test_aggregation.cpp:171-231 (tolerance 1e-6 M1)."""
    x, lx, ilr = _ref_fixture()
    N = 1e12 * np.exp(-np.arange(64) / 10.0)
    rhs, _, _ = oracle_mod.aggregation_rhs(N, x, kernel, b0, bbr, bsh, lx, ilr)
    assert abs((rhs * x).sum()) <= 1e-6 * (N * x).sum()


@pytest.mark.parametrize("sel,v_ref,alpha,v_min", [(0, 1e-18, 1, 0), (1, "mid", 1, 0),
                                                   (2, "mid", 2, 0), (3, 1.0, 1, "q")])
def test_reference_breakage_conservation(oracle_mod, sel, v_ref, alpha, v_min):
    """test_breakage.cpp:168-221 (tolerance 1e-4 M1)."""
    x, _, _ = _ref_fixture()
    v_ref = x[32] if v_ref == "mid" else v_ref
    v_min = x[16] if v_min == "q" else v_min
    N = 1e12 * np.exp(-np.arange(64) / 10.0)
    rhs, _, _ = oracle_mod.breakage_rhs(N, x, [0.5], [2.0], sel, 1e-3, v_ref, alpha, v_min)
    assert abs((rhs * x).sum()) <= 1e-4 * (N * x).sum()


def test_reference_error_codes(oracle_mod):
    x, lx, ilr = _ref_fixture()
    with pytest.raises(oracle_mod.OracleError) as e:
        oracle_mod.aggregation_rhs(np.zeros(64), x, 0, n=0)
    assert e.value.code == 1
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.breakage_rhs(np.zeros(64), x, [0.5], [2.0], 0, n=0)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.aggregation_rhs(np.zeros(64), x, 9)


# ----------------------------------------------------------- closed form --

def test_smoluchowski_closed_form_rk4(oracle_mod):
    """Constant kernel: dM0/dt = -beta0 M0^2 / 2 holds exactly for the discrete
    system, so M0(t) = M0(0) / (1 + beta0 M0(0) t / 2) up to RK4 truncation."""
    from pbe_hip import problems as P
    x, N, agg = P.config_smoluchowski(256)
    M00 = N.sum()
    Nt = oracle_mod.rk4(N, x, 1.0, 1000, agg=agg)
    exact = M00 / (1.0 + 1e-12 * M00 * 1000.0 / 2.0)
    rel = abs(Nt.sum() - exact) / exact
    assert rel <= 1e-6, rel
    # volume is conserved while nothing reaches the last pivot
    assert abs((x * Nt).sum() - (x * N).sum()) <= 1e-12 * (x * N).sum()


def test_port_matches_oracle(oracle_mod):
    from pbe_hip import problems as P
    x, N, agg, brk, t_q, bw_q = P.config_flocculation(512, daughter="uniform", n_quad=8)
    ra, Aa, _ = oracle_mod.aggregation_rhs(N, x, **agg)
    rb, Ab, _ = oracle_mod.breakage_rhs(N, x, t_q, bw_q, **brk)
    pa = oracle_mod.port_aggregation_rhs(N, x, threads=4, **agg)
    pb = oracle_mod.port_breakage_rhs(N, x, t_q, bw_q, threads=4, **brk)
    assert_parity(pa + pb, ra + rb, Aa + Ab, tol=1e-13, what="port")


@pytest.mark.parametrize("case", ["sum", "product", "breakage"])
def test_oracle_rk4_moment_laws(oracle_mod, case):
    """The oracle's RK4 against the discrete moment laws (sum kernel: M0
    decays by RK4's own polynomial; product kernel and linear-selection
    breakage: M0 linear in t; M1 conserved)."""
    import moment_cases as MC
    x, N0, kw, M0_expected, M1 = MC.CASES[case]()
    N = oracle_mod.rk4(N0, x, MC.DT, MC.STEPS, **kw)
    MC.check(x, N, M0_expected, M1, tol_m0=1e-12)


def _oracle_euler(orc, case):
    """Explicit Euler as examples/common/pbe_examples_utils.hpp:391-407:
    rhs from zero, then N_i += dt * rhs_i."""
    N = np.array(case["N0"], dtype=float)
    kw = case["kw"]
    for _ in range(case["steps"]):
        r = np.zeros_like(N)
        if "agg" in kw:
            r += orc.aggregation_rhs(N, case["x"], **kw["agg"])[0]
        if "brk" in kw:
            r += orc.breakage_rhs(N, case["x"], kw["t_q"], kw["bw_q"], **kw["brk"])[0]
        N = N + case["dt"] * r
    return N


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_oracle_reference_examples(oracle_mod, k):
    """The reference's five example programs (docs/examples.md): the oracle
    reproduces their documented errors against the analytic moments."""
    import reference_examples as RE
    c = RE.example(k)
    if c["method"] == "rk4":
        N = oracle_mod.rk4(c["N0"], c["x"], c["dt"], c["steps"], **c["kw"])
    else:
        N = _oracle_euler(oracle_mod, c)
    print(k, RE.assert_matches(c, N))
