import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ode-0.12_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def root():
    return ROOT


def golden(name):
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"), allow_pickle=False))


def say(*a):
    """Progress line naming the case (flushed before each GPU step)."""
    print(*a, flush=True)


@pytest.fixture(scope="module")
def odemi_mod():
    import torch
    assert torch.cuda.is_available(), "the GPU tests need an MI355X"
    import odemi
    return odemi
