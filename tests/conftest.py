import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-stream-join-benchmark_amd")
sys.path.insert(0, os.path.join(PKG, "py"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def sjbmod():
    import sjb
    sjb.lib()
    return sjb


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(sjbmod):
    c = sjbmod.Context(device=0, max_window=1 << 22, distinct_keys=1 << 16)
    yield c
    c.close()


@pytest.fixture(scope="session")
def torch_mod():
    import torch
    return torch


def to_dev(torch, arr):
    """numpy tuple array -> device buffer (torch uint8 tensor)."""
    if len(arr) == 0:
        return torch.empty(16, dtype=torch.uint8, device="cuda")
    t = torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy())
    return t.to("cuda")
