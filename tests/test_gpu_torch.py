"""torch.autograd wrapper over the RK4 adjoint (own process: torch loads the
HIP runtime first)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_torch_autograd_gradient_and_fit():
    try:
        import importlib.util
        if importlib.util.find_spec("torch") is None:
            pytest.skip("torch not installed")
    except Exception:  # noqa: BLE001
        pytest.skip("torch not importable")
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "torch_fit_check.py")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    g, fd = np.array(out["grad"]), np.array(out["fd"])
    assert np.all(np.abs(g - fd) <= 1e-5 * np.abs(fd) + 1e-9 * out["loss"]), (g, fd)
    assert out["hist"][-1] < out["hist"][0]
