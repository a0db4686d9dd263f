import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "pbe-fixed-pivot-cuda_amd", "python"))
sys.path.insert(0, os.path.join(REPO, "pbe-fixed-pivot-cuda_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (run on the MI355X box)")


def pytest_collection_modifyitems(config, items):
    has_gpu = None
    for item in items:
        if "gpu" in item.keywords:
            if has_gpu is None:
                from pbe_hip import hip
                has_gpu = hip.device_count() > 0
            if not has_gpu:
                item.add_marker(pytest.mark.skip(reason="no HIP device visible"))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def pbe():
    import pbe_hip
    pbe_hip.library()
    return pbe_hip
