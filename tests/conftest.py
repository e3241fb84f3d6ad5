import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import hitomi_amd

    c = hitomi_amd.Context(0)
    yield c
    c.close()


def uniform_catalogue(n, L, seed, rsd_sigma=0.0):
    """Uniform positions in [0, L)^3, optionally shifted along z by N(0, sigma) and wrapped
    once (particleGadget.hpp:565-591)."""
    rng = np.random.default_rng(seed)
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    pos = rng.random((n, 3)) * L
    if rsd_sigma > 0:
        pos[:, 2] += rng.normal(0.0, rsd_sigma, n)
        hi = pos[:, 2] >= L[2]
        lo = pos[:, 2] < 0.0
        pos[hi, 2] -= L[2]
        pos[lo, 2] += L[2]
    return pos


def clustered_catalogue(n, L, seed, nclump=40, sigma=20.0):
    """Gaussian clumps plus a uniform background, wrapped into [0, L)^3."""
    rng = np.random.default_rng(seed)
    L = float(L)
    centres = rng.random((nclump, 3)) * L
    k = n // 2
    idx = rng.integers(0, nclump, k)
    a = centres[idx] + rng.normal(0.0, sigma, (k, 3))
    b = rng.random((n - k, 3)) * L
    pos = np.concatenate([a, b])
    pos = np.mod(pos, L)
    pos[pos >= L] -= L
    return pos


def report_err(name, value):
    """Print a measured error next to its tolerance check, so a run with -s lists
    the margins of every FP32 comparison."""
    test = os.environ.get("PYTEST_CURRENT_TEST", "").split(" ")[0]
    print(f"\nTOLREPORT {test} {name}: {float(value):.3e}", flush=True)


def rel_check(name, got, ref, tol, scale=None):
    """max |got - ref| / scale (default max |ref|), reported, then checked against tol."""
    scale = np.abs(ref).max() if scale is None else scale
    e = float(np.abs(np.asarray(got) - np.asarray(ref)).max() / scale)
    report_err(name, e)
    assert e <= tol, (name, e, tol)
    return e
