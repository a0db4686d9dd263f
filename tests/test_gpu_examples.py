"""The example programs (pbe-fixed-pivot-cuda_amd/examples), the counterparts
of the reference's examples/, run on the device: each prints its errors
against the analytic moment laws as a JSON line, which must match what the
reference documents for it (docs/examples.md, restated with the exact
bounds in tests/reference_examples.py).  The two-stage flocculation program
is checked against the oracle's RK4: a 60-step run (shear switched after 30)
entry by entry, and the full 3600-step run against the committed oracle
fixture (d43 at t_mid and t_end, M0, the first-moment drift, the final
state)."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "examples")
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))


def _run(name, *args, timeout=110):
    exe = os.path.join(EX, name, "example_" + name)
    assert os.path.exists(exe), f"{exe} not built (run build())"
    r = subprocess.run([exe, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    print(r.stdout[-2500:])
    assert r.returncode == 0, r.stdout[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("k,name", [(1, "aggregation_constant"), (2, "aggregation_sum"),
                                    (3, "aggregation_product"),
                                    (4, "breakage_linear_symmetricbinary"),
                                    (5, "breakage_linear_uniform")])
def test_example_program_matches_documented_errors(k, name):
    import reference_examples as RE
    case = RE.example(k)
    out = _run(name)
    for key, fig in case["documented"].items():
        assert abs(out[key] - fig) <= 0.05 * fig, (key, out[key], fig)
    for key, bound in case["exact"].items():
        assert out[key] <= bound, (key, out[key], bound)


def test_flocculation_prefix_vs_oracle(oracle_mod, tmp_path):
    """N = 960, 60 RK4 steps of 0.5 s with the shear rate switched after the
    step starting at t = 15 s: the program's final state against the
    oracle's RK4 on the same schedule."""
    from pbe_hip import problems as P
    c = P.config_flocculation_two_stage(960, t_end_s=30.0, t_mid_s=15.0)
    f = tmp_path / "state.bin"
    out = _run("flocculation_two_stage", 960, 30, 15, f)
    got = np.fromfile(f, dtype=np.float64)
    assert out["steps"] == c["steps"] == 60 and out["steps_high"] == c["n_high"]
    kw = dict(t_q=c["t_q"], bw_q=c["bw_q"], threads=ORACLE_THREADS)
    ref = oracle_mod.rk4(c["N0"], c["x"], c["dt"], c["n_high"], agg=c["agg_high"],
                         brk=c["brk_high"], **kw)
    ref = oracle_mod.rk4(ref, c["x"], c["dt"], c["steps"] - c["n_high"], agg=c["agg_low"],
                         brk=c["brk_low"], **kw)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("flocculation 60 steps: max |gpu - oracle| / max|oracle| =", err)
    assert err <= 1e-11
    assert abs(out["d43_end_um"] - P.d43(ref, c["x"], c["d0"]) * 1e6) <= 1e-11 * out["d43_end_um"]


def test_flocculation_full_run_vs_oracle_fixture(tmp_path):
    """The case study as the reference runs it (N = 960, 3600 RK4 steps,
    G = 100 -> 20 1/s): d43 at t_mid and t_end, M0 and the final state
    against the oracle's run (tests/golden/flocculation_two_stage.json); the
    first moment is conserved (erosion keeps volume; nothing reaches the
    last section)."""
    with open(os.path.join(REPO, "tests", "golden", "flocculation_two_stage.json")) as fh:
        fx = json.load(fh)
    f = tmp_path / "state.bin"
    out = _run("flocculation_two_stage", 960, 1800, 600, f)
    got = np.fromfile(f, dtype=np.float64)
    ref = np.array(fx["N_end"])
    assert out["steps"] == fx["steps"] and out["steps_high"] == fx["steps_high"]
    for key in ("d43_mid_um", "d43_end_um", "M0_end"):
        assert abs(out[key] - fx[key]) <= 1e-9 * abs(fx[key]), (key, out[key], fx[key])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("flocculation 3600 steps: max |gpu - oracle| / max|oracle| =", err,
          "wall", out["wall_s"], "s")
    assert err <= 1e-9
    assert out["M1_drift"] <= 1e-10 and fx["M1_drift"] <= 1e-10
