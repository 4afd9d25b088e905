import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "oracle"), os.path.join(REPO, "nextla.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")

# Tolerances of the reference's test suite (test/lapack_helpers.jl:14).
RTOL = {np.float64: 1e-12, np.float32: 1e-5}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def rtol_for(dtype):
    return RTOL[np.dtype(dtype).type]


def _wide(x):
    # complex inputs keep their imaginary parts
    x = np.asarray(x)
    return x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)


def isapprox(x, y, rtol):
    """Julia's isapprox for arrays: norm(x - y) <= rtol * max(norm(x), norm(y))."""
    x, y = _wide(x), _wide(y)
    d = np.linalg.norm(x - y)
    return d <= rtol * max(np.linalg.norm(x), np.linalg.norm(y))


def relerr(x, y):
    x, y = _wide(x), _wide(y)
    den = max(np.linalg.norm(y), 1e-300)
    return np.linalg.norm(x - y) / den


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def load(name):
        if name not in cache:
            cache[name] = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
        return cache[name]

    return load
