import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libvesicle.so")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def vb():
    import cudabubble_amd as vb
    vb.lib()
    return vb


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when no GPU is visible: GPU tests must run the HIP path."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch.device("cuda:0")


def grid(n, nz=None, L=2 * np.pi):
    nz = n if nz is None else nz
    x = np.arange(n) * L / n - L / 2          # fluidbase.cpp:156-161
    z = np.arange(nz) * L / nz - L / 2
    Z, Y, X = np.meshgrid(z, x, x, indexing="ij")
    return X, Y, Z


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def flush_print(*a):
    print(*a, flush=True)
