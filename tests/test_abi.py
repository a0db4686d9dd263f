"""CPU tests of the C-ABI library: it loads without a GPU, exports every
function declared in include/pbe_hip/*.h, its parameter structs have the
reference's layout, and every argument check returns the reference's error
code before any HIP runtime call (src/aggregation.cu:23-24,54;
src/breakage.cu:23-24,46,74)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(REPO, "include", "pbe_hip")

INVALID_VALUE = 1
INVALID_CONFIGURATION = 9
INVALID_DEVICE_POINTER = 17


def _declared_functions():
    names = set()
    for fn in os.listdir(HEADER_DIR):
        if fn.endswith(".h"):
            text = open(os.path.join(HEADER_DIR, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(pbe_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol(pbe):
    lib = pbe.library()
    names = _declared_functions()
    assert "pbe_launch_aggregation_rhs" in names and "pbe_launch_breakage_rhs" in names
    for name in names:
        assert hasattr(lib, name), f"{name} declared but not exported"


def test_param_struct_layouts(pbe):
    A, B = pbe.AggregationParams, pbe.BreakageParams
    assert ctypes.sizeof(A) == 64 and ctypes.sizeof(B) == 56
    assert [getattr(A, f).offset for f, _ in A._fields_] == [0, 8, 16, 24, 32, 40, 48, 56]
    assert [getattr(B, f).offset for f, _ in B._fields_] == [0, 4, 8, 16, 24, 32, 40, 48]
    a = A()
    assert (a.n, a.log_x0, a.inv_log_r, a.kernel_type, a.beta0, a.beta_br, a.beta_sh,
            a.block_size) == (0, 0.0, 0.0, 0, 1.0, 0.0, 0.0, 256)
    b = B()
    assert (b.n, b.n_quad, b.selection, b.S0, b.v_ref, b.alpha, b.v_min,
            b.block_size) == (0, 0, 0, 1.0, 1.0, 1.0, 0.0, 256)


FAKE = 0x1000  # never dereferenced: every case below fails validation first


@pytest.mark.parametrize("kw,ptrs,code", [
    (dict(n=0), (FAKE, FAKE, FAKE), INVALID_VALUE),
    (dict(n=-5), (FAKE, FAKE, FAKE), INVALID_VALUE),
    (dict(n=0), (None, None, None), INVALID_VALUE),            # n is checked first
    (dict(n=10), (None, FAKE, FAKE), INVALID_DEVICE_POINTER),
    (dict(n=10), (FAKE, None, FAKE), INVALID_DEVICE_POINTER),
    (dict(n=10), (FAKE, FAKE, None), INVALID_DEVICE_POINTER),
    (dict(n=10, kernel_type=6), (FAKE, FAKE, FAKE), INVALID_VALUE),
    (dict(n=10, kernel_type=-1), (FAKE, FAKE, FAKE), INVALID_VALUE),
    (dict(n=10, kernel_type=6), (None, FAKE, FAKE), INVALID_DEVICE_POINTER),  # ptr before enum
    (dict(n=10, block_size=2048), (FAKE, FAKE, FAKE), INVALID_CONFIGURATION),
])
def test_aggregation_argument_errors(pbe, kw, ptrs, code):
    p = pbe.AggregationParams(**kw)
    assert pbe.launch_aggregation_rhs(*ptrs, p) == code


@pytest.mark.parametrize("kw,ptrs,code", [
    (dict(n=0, n_quad=1), (FAKE,) * 5, INVALID_VALUE),
    (dict(n=10, n_quad=0), (FAKE,) * 5, INVALID_VALUE),
    (dict(n=10, n_quad=-1), (None,) * 5, INVALID_VALUE),
    (dict(n=10, n_quad=1), (None, FAKE, FAKE, FAKE, FAKE), INVALID_DEVICE_POINTER),
    (dict(n=10, n_quad=1), (FAKE, FAKE, None, FAKE, FAKE), INVALID_DEVICE_POINTER),
    (dict(n=10, n_quad=1), (FAKE, FAKE, FAKE, None, FAKE), INVALID_DEVICE_POINTER),
    (dict(n=10, n_quad=1), (FAKE, FAKE, FAKE, FAKE, None), INVALID_DEVICE_POINTER),
    (dict(n=10, n_quad=1, selection=4), (FAKE,) * 5, INVALID_VALUE),
    (dict(n=10, n_quad=1, block_size=0), (FAKE,) * 5, INVALID_VALUE),
    (dict(n=10, n_quad=1, block_size=4096), (FAKE,) * 5, INVALID_CONFIGURATION),
])
def test_breakage_argument_errors(pbe, kw, ptrs, code):
    p = pbe.BreakageParams(**kw)
    assert pbe.launch_breakage_rhs(*ptrs, p) == code


def test_null_params_pointer(pbe):
    lib = pbe.library()
    assert lib.pbe_launch_aggregation_rhs(FAKE, FAKE, FAKE, None, None) == INVALID_VALUE
    assert lib.pbe_launch_breakage_rhs(FAKE, FAKE, FAKE, FAKE, FAKE, None, None) == INVALID_VALUE


def test_version(pbe):
    assert "gfx950" in pbe.version()


def test_daughter_tables():
    """Moments of the daughter tables: number sum bw and volume sum bw t."""
    import numpy as np
    from pbe_hip import problems as P
    t, bw = P.daughter_symmetric()
    assert bw.sum() == 2.0 and (bw * t).sum() == 1.0
    t, bw = P.daughter_erosion(0.05)
    assert bw.sum() == 2.0 and abs((bw * t).sum() - 1.0) < 1e-15
    for m in (4, 8, 16, 32):
        t, bw = P.daughter_uniform(m)
        assert abs(bw.sum() - 2.0) < 1e-14 and abs((bw * t).sum() - 1.0) < 1e-14
        assert np.all((t > 0) & (t < 1))
    for q in (0.0, 1.0, 2.0, 3.0):
        for m in (8, 16, 32):
            t, bw = P.daughter_power_law(q, m)
            # b = (q+2)/v' (v/v')^q integrates to nu = (q+2)/(q+1) fragments and
            # conserves volume; GL with m/2 points per panel is exact for
            # integer q < m - 1
            assert abs(bw.sum() - (q + 2) / (q + 1)) < 1e-13
            assert abs((bw * t).sum() - 1.0) < 1e-13
    with pytest.raises(ValueError):
        P.daughter_power_law(2.0, 7)


def test_birth_pairs_brute_force():
    """problems.birth_pairs (the pairs the roofline charges: fl(x_j + x_k) !=
    x_k) against a pair-by-pair count on geometric grids wide enough that
    most pairs have a zero birth weight, a jittered grid, and the edge of a
    binade (x_j exactly half an ulp of x_k rounds to even)."""
    import numpy as np
    from pbe_hip import problems as P

    def brute(x):
        return sum(1 for k in range(len(x)) for j in range(k + 1) if x[j] + x[k] != x[k])

    for x in (P.geometric_grid(500, 1e-18, 1.1), P.geometric_grid(300, 1e-18, 2 ** 0.25),
              P.jittered_grid(400, 1e-18, 1.12, 0.2, seed=5)):
        assert P.birth_pairs(x) == brute(x)
    # 1 + 2^-53 rounds to 1 (tie to even), 1.5 ulp above 1 + 2^-52 does not
    x = np.array([2.0 ** -54, 2.0 ** -53, 2.0 ** -52, 1.0, 1.0 + 2.0 ** -52])
    assert P.birth_pairs(x) == brute(x)
    assert P.birth_pairs(P.geometric_grid(16384, 1e-18, 1.02)) < 16384 * 16385 // 4


def test_device_cube_root_host_check():
    """cbrt_pair (the bit-guess + Newton cube root of the kernels), compiled
    for the host from the same header: within 3 ulp of the C library's cbrt,
    its reciprocal within 6 ulp (tests/cpp/test_roots.cpp)."""
    import json
    import subprocess
    exe = os.path.join(REPO, "tests", "cpp", "test_roots")
    assert os.path.exists(exe), "tests/cpp/test_roots not built (run build())"
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=120)
    out = json.loads(r.stdout)
    assert r.returncode == 0, out
    assert out["cbrt_max_ulp"] <= 4 and out["inv_cbrt_max_ulp"] <= 8


def test_strip_plan_host_check():
    """The strip launch's work split (agg_strip.hpp), compiled for the host:
    every chunk of every column block goes to exactly one workgroup share,
    the cost line equals the sum of its chunk costs, and the waves' row
    ranges tile each visit (tests/cpp/test_strip_plan.cpp)."""
    import json
    import subprocess
    exe = os.path.join(REPO, "tests", "cpp", "test_strip_plan")
    assert os.path.exists(exe), "tests/cpp/test_strip_plan not built (run build())"
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=120)
    out = json.loads(r.stdout)
    assert r.returncode == 0 and out["failures"] == 0, out
    assert out["chunks"] > 1e7


def test_deterministic_argument_errors(pbe):
    """The gather-mode launches check arguments, then the workspace size,
    before any HIP call."""
    lib = pbe.library()
    A = pbe.AggregationParams(n=100, kernel_type=5)
    B = pbe.BreakageParams(n=100, n_quad=1)
    need_a = lib.pbe_aggregation_deterministic_workspace_bytes(100)
    need_b = lib.pbe_breakage_deterministic_workspace_bytes(100)
    assert need_a > 0 and need_b > 0
    assert lib.pbe_aggregation_deterministic_workspace_bytes(0) == 0
    agg = lib.pbe_launch_aggregation_rhs_deterministic
    brk = lib.pbe_launch_breakage_rhs_deterministic
    assert agg(FAKE, FAKE, FAKE, None, FAKE, need_a, None) == INVALID_VALUE
    assert agg(FAKE, FAKE, FAKE, ctypes.byref(A), None, need_a, None) == INVALID_DEVICE_POINTER
    assert agg(FAKE, FAKE, FAKE, ctypes.byref(A), FAKE, need_a - 8, None) == INVALID_VALUE
    bad = pbe.AggregationParams(n=100, kernel_type=7)
    assert agg(FAKE, FAKE, FAKE, ctypes.byref(bad), FAKE, need_a, None) == INVALID_VALUE
    assert brk(FAKE, FAKE, FAKE, FAKE, FAKE, ctypes.byref(B), None, need_b,
               None) == INVALID_DEVICE_POINTER
    assert brk(FAKE, FAKE, FAKE, FAKE, FAKE, ctypes.byref(B), FAKE, need_b - 8,
               None) == INVALID_VALUE
    zero = pbe.BreakageParams(n=100, n_quad=0)
    assert brk(FAKE, FAKE, FAKE, FAKE, FAKE, ctypes.byref(zero), FAKE, need_b,
               None) == INVALID_VALUE


def test_aggregation_option_errors(pbe):
    """tile_u outside {0,1,2,4,8}, band_split outside {0,1,2,4,8,16} or form
    outside {0,1,2,3,4} are refused."""
    lib = pbe.library()
    A = pbe.AggregationParams(n=100, kernel_type=5)
    for u, w, f in ((3, 0, 0), (16, 0, 0), (4, 3, 0), (4, 32, 0), (4, -1, 0), (4, 4, 5),
                    (4, 4, -1)):
        opt = pbe.AggOptions(u, w, f)
        assert lib.pbe_launch_aggregation_rhs_ex(FAKE, FAKE, FAKE, ctypes.byref(A),
                                                 ctypes.byref(opt), None) == INVALID_VALUE


def _table_rows(name, width):
    path = os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "csrc", name)
    rows = [tuple(int(v) for v in m.group(1).split(","))
            for m in re.finditer(r"^\{([-0-9, ]+)\},$", open(path).read(), flags=re.M)]
    assert rows and all(len(r) == width for r in rows)
    return rows


@pytest.mark.parametrize("name,width", [("tile_table.inc", 6), ("batch_tile_table.inc", 4)])
def test_tile_tables_cover_every_case(name, width):
    """Every kernel and every grid (without the geometric lookup, and every
    reach) finds a row at every n: per (kernel, reach class) the rows' n_max
    increase and the last is open, and each kernel has a row of reach -1 and
    one of open reach.  Shapes are ones the launches take."""
    rows = _table_rows(name, width)
    groups = {}
    for r in rows:
        assert r[1] in range(-1, 6) and r[3] in (1, 2, 4, 8)
        if width == 6:
            assert r[4] in (1, 2, 4, 8, 16) and r[5] in (1, 2, 3, 4)
        groups.setdefault((r[1], r[2]), []).append(r[0])
    for key, nmax in groups.items():
        assert nmax == sorted(nmax) and len(set(nmax)) == len(nmax), key
        assert nmax[-1] == 1 << 30, key
    for k in range(6):
        reaches = {c for (kk, c) in groups if kk in (k, -1)}
        assert -1 in reaches and (1 << 30) in reaches, (name, k)


def test_default_options_follow_the_table(pbe):
    """pbe_aggregation_default_options resolves (n, kernel, reach) through
    the first matching row of tile_table.inc; bad arguments are refused."""
    import math
    rows = _table_rows("tile_table.inc", 6)
    for n, r, k in ((300, 1.02, 5), (5000, 1.01, 3), (70000, 1.002, 0), (2048, None, 2),
                    (1000, 2.0, 1)):
        il = 0.0 if r is None else 1.0 / math.log(r)
        reach = -1.0 if r is None else max(0.0, math.log(1.0 / (r - 1.0)) / math.log(r))
        want = next((u, w, f) for (nm, kk, rm, u, w, f) in rows
                    if n <= nm and kk in (k, -1)
                    and (rm < 0 if reach < 0 else (rm >= 0 and reach <= rm)))
        got = pbe.aggregation_default_options(pbe.AggregationParams(n=n, kernel_type=k,
                                                                    inv_log_r=il))
        assert got == want, (n, r, k)
    lib = pbe.library()
    opt = pbe.AggOptions(0, 0, 0)
    for A in (pbe.AggregationParams(n=0), pbe.AggregationParams(n=10, kernel_type=6)):
        assert lib.pbe_aggregation_default_options(ctypes.byref(A),
                                                   ctypes.byref(opt)) == INVALID_VALUE
    assert lib.pbe_aggregation_default_options(None, ctypes.byref(opt)) == INVALID_VALUE


def test_batch_plan_tile_width_refused():
    """pbe_batch_plan_create_ex refuses a tile width other than 0, 1, 2, 4, 8
    before any device work."""
    import pbe_hip as pbe
    inst = (pbe.BatchInstance * 1)()
    inst[0].n = 64
    h = ctypes.c_void_p()
    for u in (3, 16, -1):
        assert pbe.library().pbe_batch_plan_create_ex(ctypes.byref(h), inst, 1, 5, -1, 0,
                                                      u) == INVALID_VALUE
        assert not h.value


def test_torch_solve_rejects_host_tensors():
    """PBESolve refuses a host tensor before any device work (its pointer
    would otherwise reach the kernels)."""
    torch = pytest.importorskip("torch")
    from pbe_hip import problems as P
    from pbe_hip.torch_ops import PBESolve, Problem
    x, N0, agg, brk, t_q, bw_q = P.config_adjoint(64)
    prob = Problem(x, agg, brk, t_q, bw_q, 0.5, 2)
    theta = torch.tensor([agg["beta_br"], agg["beta_sh"], brk["S0"], brk["alpha"]],
                         dtype=torch.float64)
    with pytest.raises(ValueError, match="HIP device"):
        PBESolve.apply(torch.tensor(N0), theta, prob)


def test_noalloc_argument_errors(pbe):
    """The no-allocation gather checks its arguments like the workspace form,
    and refuses n beyond its static array, before any HIP call."""
    lib = pbe.library()
    f = lib.pbe_launch_aggregation_rhs_deterministic_noalloc
    assert f(FAKE, FAKE, FAKE, None, None) == INVALID_VALUE
    assert f(FAKE, FAKE, FAKE, ctypes.byref(pbe.AggregationParams(n=0)), None) == INVALID_VALUE
    assert f(None, FAKE, FAKE, ctypes.byref(pbe.AggregationParams(n=10)), None) == \
        INVALID_DEVICE_POINTER
    assert f(FAKE, FAKE, FAKE, ctypes.byref(pbe.AggregationParams(n=10, kernel_type=7)),
             None) == INVALID_VALUE
    assert f(FAKE, FAKE, FAKE, ctypes.byref(pbe.AggregationParams(n=65537)), None) == \
        INVALID_VALUE


def test_flocculation_two_stage_setup():
    """The case study's dimensionless set-up (reference
    examples/flocculation_two_stage/main.cpp): the log-normal start has
    M0 = 1 and M1 = 1.1445 on its grid (the figure the reference's utilities
    quote, examples/common/pbe_examples_utils.hpp:141); the shear rate
    switches after the step that starts below t = 600 s by the accumulated
    dimensionless time, which is step 1201 of 3600; the dimensionless shear
    and selection coefficients follow from the physical constants."""
    import math
    from pbe_hip import problems as P
    c = P.config_flocculation_two_stage(960)
    assert abs(c["N0"].sum() - 1.0) < 1e-15
    assert round(float((c["x"] * c["N0"]).sum()), 4) == 1.1445
    assert c["steps"] == 3600 and c["n_high"] == 1201 and c["mid_step"] == 1200
    assert abs(c["agg_high"]["beta_sh"] / c["agg_low"]["beta_sh"] - 5.0) < 1e-14
    assert abs(c["t_c"] - 119.30821604109944) < 1e-9
    d = P.FLOC_CONSTANTS
    assert abs(c["x"][0] - (d["d_min"] / d["d0"]) ** 3) < 1e-18
    assert abs(c["x"][-1] / (d["d_max"] / d["d0"]) ** 3 - 1.0) < 1e-10
    assert c["brk_high"]["S0"] == d["k_br"] * 100.0 * c["t_c"]
    assert math.isclose(P.d43(c["N0"], c["x"], 1.0), float((c["x"] ** (4 / 3) * c["N0"]).sum()
                                                            / (c["x"] * c["N0"]).sum()))
