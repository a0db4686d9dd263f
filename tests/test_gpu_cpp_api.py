"""The reference's 31 unit tests (tests/unit/test_aggregation.cpp and
tests/unit/test_breakage.cpp), restated in C++ against the drop-in header
<pbe_cuda/pbe_cuda.cuh> and run as one program on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "tests", "cpp", "test_pbe_api")


def test_reference_unit_tests_cpp_api():
    assert os.path.exists(EXE), "tests/cpp/test_pbe_api not built (run build())"
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "31 tests, 0 failed" in r.stdout
