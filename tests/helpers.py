"""Shared helpers for the test suite."""
import numpy as np

RTOL = 1e-10  # FP64 parity bar: max |a-b| / max |b|


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-300)
    return float(np.abs(a - b).max()) / scale if b.size else 0.0


def assert_close(a, b, tol=RTOL, what=""):
    r = rel(a, b)
    assert r <= tol, f"{what}: relative error {r:.3e} > {tol:.1e}"
