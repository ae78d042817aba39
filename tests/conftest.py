import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_body(name):
    """Body file rows "x z" (src/VortexSheet.cpp:56-76 reads them as raw ASCII)."""
    a = np.loadtxt(os.path.join(GOLDEN, name))
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1])


def config2_blobs(n, seed=2024):
    """Isolated-sum workload: x~U[-1,15], z~U[-3,3], circ~N(0, 2e-4) mean
    removed, sigma = 0.01."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 15.0, n)
    z = rng.uniform(-3.0, 3.0, n)
    c = rng.normal(0.0, 2e-4, n)
    c -= c.mean()
    s = np.full(n, 0.01)
    return x, z, c, s


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    import dvmpp_amd
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    ctx = dvmpp_amd.Context(0)
    yield ctx
    ctx.close()


def to_dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda") for a in arrs]


def rel_err(a, b):
    """max |a-b| / max |b| over a vector (normwise-max relative error)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)
