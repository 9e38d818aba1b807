"""Tolerances of the parity tests.

complex128 results: |d| <= 1e-10 * max(|ref|, 1e-3 * ||ref||_inf), relative
with an absolute floor for entries near zero.  complex64 results: within
1e-5 of the complex128 oracle, scaled by max(1, ||ref||_inf).
"""
import numpy as np


def close128(got, ref, rel=1e-10):
    got, ref = np.asarray(got), np.asarray(ref)
    scale = np.maximum(np.abs(ref), 1e-3 * max(np.max(np.abs(ref)) if ref.size else 0.0, 1e-300))
    err = np.abs(got - ref)
    assert np.all(err <= rel * scale + 1e-300), f"max rel err {np.max(err / scale):.3e}"


def close64(got, ref, tol=1e-5):
    got, ref = np.asarray(got), np.asarray(ref)
    assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref))), \
        f"max abs err {np.max(np.abs(got - ref)):.3e}"
