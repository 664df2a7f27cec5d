import os
import sys

import numpy as np
import pytest

# torch before the engine library: torch and libcgfd3d.so both bind
# libamdhip64.so.7, the first one loaded serves the whole process, and torch
# initialises only on the runtime it ships with, so the surrogate's GPU tests
# fail in a process that loaded the engine first
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("copara-cgfd3d_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcgfd3d.so")
    config.addinivalue_line("markers", "slow: long-running case")


def assert_same(got, want, what=""):
    """Bit-for-bit value equality (+0 and -0 compare equal, NaN never does)."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} vs {want.shape}"
    assert np.isfinite(want).all(), f"{what}: the reference values are not finite"
    bad = ~(got == want)
    if bad.any():
        idx = np.argwhere(bad)
        first = tuple(idx[0])
        diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
        scale = np.abs(want.astype(np.float64)).max() or 1.0
        raise AssertionError(f"{what}: {bad.sum()} of {bad.size} values differ; first at {first}: "
                             f"{got[first]!r} vs {want[first]!r}; max |diff| {np.nanmax(diff):.3e} "
                             f"(max |ref| {scale:.3e})")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum(a * a)))
