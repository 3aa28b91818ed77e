"""Shared test setup.

Tests marked ``gpu`` need an MI355X (they call libsexgpu.so through its C-ABI); the
others run on the CPU (oracle against golden vectors, host logic, library exports).
"""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "gpu-sextractor_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _gpu_available():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def masks():
    d = np.load(os.path.join(GOLDEN, "masks.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def frame1024():
    import sexgpu
    return sexgpu.synth("plain1024")
