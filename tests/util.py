"""Shared helpers of the test suite: seeded operands and error bounds.

Tolerances follow the forward-error bound of a length-K inner product,
|fl(s) - s| <= gamma_K * sum|terms|, gamma_K ~ K*eps, with a small constant
factor for the different association orders of the GPU reduction trees and the
oracle's sequential loops.  Elementwise kernels (one multiply-add per element)
are held to a few ulps of the terms' magnitude.
"""
from __future__ import annotations

import numpy as np

EPS = {np.dtype(np.float32): float(np.finfo(np.float32).eps), np.dtype(np.float64): float(np.finfo(np.float64).eps)}


def eps(dtype) -> float:
    return EPS[np.dtype(dtype)]


def uniform(rng, shape, dtype, lo=-1.0, hi=1.0):
    return rng.uniform(lo, hi, shape).astype(dtype)


def vec(rng, n, inc, dtype):
    """Padded buffer for an n-vector at stride |inc| (reference AlignedBuffer
    layout, test/common/generate_buffer.h:57-68), U[-1,1] everywhere."""
    size = 1 + (n - 1) * abs(inc) if n > 0 else 1
    return uniform(rng, size, dtype)


def elems(buf, n, inc):
    """The n logical elements of a padded buffer, element 0 first."""
    if n == 0:
        return buf[:0]
    idx = np.arange(n) * abs(inc)
    if inc < 0:
        idx = idx[::-1]
    return buf[idx]


def assert_bound(got, want, mag, K, dtype, c=4.0, what=""):
    """|got - want| <= c * (K + 2) * eps * mag + tiny, elementwise."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    mag = np.asarray(mag, dtype=np.float64)
    tol = c * (K + 2) * eps(dtype) * mag + 1e-30
    err = np.abs(got - want)
    bad = ~(err <= tol)
    if np.any(bad):
        i = int(np.flatnonzero(bad.ravel())[0])
        raise AssertionError(
            f"{what}: {int(bad.sum())} element(s) out of bound; first at flat index {i}: got "
            f"{got.ravel()[i]!r} want {want.ravel()[i]!r} err {err.ravel()[i]:.3e} tol {tol.ravel()[i]:.3e}")


def max_rel_err(got, truth):
    got = np.asarray(got, dtype=np.float64)
    truth = np.asarray(truth, dtype=np.float64)
    den = np.maximum(np.abs(truth), 1e-300)
    return float(np.max(np.abs(got - truth) / den)) if got.size else 0.0


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()
