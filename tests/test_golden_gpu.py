"""The HIP path on the committed golden fixtures (tests/golden/*.npz) through
the exo_* C-ABI, against the stored oracle outputs within the K*eps bound."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import LEFT, LOWER, NO_TRANS, NON_UNIT, TRANS, UNIT, UPPER
from util import assert_bound

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RM = 101


@pytest.fixture(scope="module")
def C():
    from exoblas_amd import cblas, lib
    lib.load()
    return cblas


def load(name, dt):
    return np.load(os.path.join(HERE, f"{name}_{np.dtype(dt).name}.npz"))


def fn(C, dt, name):
    return getattr(C, ("exo_s" if dt == np.float32 else "exo_d") + name)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_level1(C, dt):
    g = load("level1", dt)
    x, y = g["x"], g["y"]
    n = len(x)
    x64, y64 = x.astype(np.float64), y.astype(np.float64)
    assert_bound(fn(C, dt, "asum")(n, x.copy(), 1), g["asum"][0], np.sum(np.abs(x64)), n, dt)
    d = fn(C, dt, "dot")(n, x.copy(), 1, y.copy(), 1)
    assert_bound(d, g["dot"][0], np.sum(np.abs(x64 * y64)), n, dt)
    assert_bound(d, g["dot_truth"][0], np.sum(np.abs(x64 * y64)), n, dt)
    yy = y.copy(); fn(C, dt, "axpy")(n, 0.75, x.copy(), 1, yy, 1)
    assert_bound(yy, g["axpy"], np.abs(y64) + 0.75 * np.abs(x64), 1, dt)
    xx, yy = x.copy(), y.copy(); fn(C, dt, "rot")(n, xx, 1, yy, 1, 0.6, 0.8)
    assert_bound(xx, g["rot_x"], np.abs(x64) + np.abs(y64), 2, dt)
    assert_bound(yy, g["rot_y"], np.abs(x64) + np.abs(y64), 2, dt)
    P = np.array([-1.0, 0.3, -0.7, 1.1, 0.45], dt)  # H = {P1, P3, P2, P4} = stored {0.3, 1.1, -0.7, 0.45}
    xx, yy = x.copy(), y.copy(); fn(C, dt, "rotm")(n, xx, 1, yy, 1, P)
    assert_bound(xx, g["rotm_x"], 2 * (np.abs(x64) + np.abs(y64)), 2, dt)
    assert_bound(yy, g["rotm_y"], 2 * (np.abs(x64) + np.abs(y64)), 2, dt)
    xx = x.copy(); fn(C, dt, "scal")(n, 2.5, xx, 1)
    assert np.array_equal(xx, g["scal"])


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_level2(C, dt):
    g = load("gemv", dt)
    A, x, y = g["A"], g["x"], g["y"]
    m, n = A.shape
    yy = y.copy(); fn(C, dt, "gemv")(RM, NO_TRANS, m, n, 2.4, A, n, x.copy(), 1, 2.4, yy, 1)
    assert_bound(yy, g["gemv_n"], 2.4 * np.abs(y) + 2.4 * (np.abs(A).astype(np.float64) @ np.abs(x)), n, dt)
    yt = g["yt"].copy(); fn(C, dt, "gemv")(RM, TRANS, m, n, 2.4, A, n, g["xt"].copy(), 1, 2.4, yt, 1)
    assert_bound(yt, g["gemv_t"], 2.4 * np.abs(g["yt"]) + 2.4 * (np.abs(A).astype(np.float64).T @ np.abs(g["xt"])),
                 m, dt)
    t = load("level2_tri", dt)
    A, D, x, y = t["A"], t["D"], t["x"], t["y"]
    n = len(x)
    absA = np.abs(A).astype(np.float64)
    yy = y.copy(); fn(C, dt, "symv")(RM, LOWER, n, 1.5, A, n, x.copy(), 1, 0.5, yy, 1)
    assert_bound(yy, t["symv_l"], 0.5 * np.abs(y) + 1.5 * ((absA + absA.T) @ np.abs(x)), n, dt)
    AA = A.copy(); fn(C, dt, "syr2")(RM, UPPER, n, 1.2, x.copy(), 1, y.copy(), 1, AA, n)
    assert_bound(AA, t["syr2_u"], np.abs(A) + 3, 2, dt)
    xx = x.copy(); fn(C, dt, "trmv")(RM, UPPER, TRANS, UNIT, n, A, n, xx, 1)
    assert_bound(xx, t["trmv_utu"], absA.T @ np.abs(x) + 2 * np.abs(x), n, dt)
    xx = x.copy(); fn(C, dt, "trsv")(RM, LOWER, NO_TRANS, NON_UNIT, n, D, n, xx, 1)
    assert_bound(xx, t["trsv_lnn"], np.abs(t["trsv_lnn"]) + 1, n, dt)
    xx = x.copy(); fn(C, dt, "trsv")(RM, UPPER, TRANS, NON_UNIT, n, D, n, xx, 1)
    assert_bound(xx, t["trsv_utn"], np.abs(t["trsv_utn"]) + 1, n, dt)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_level3(C, dt):
    g = load("level3", dt)
    A, B, Cm, S = g["A"], g["B"], g["C"], g["S"]
    M, K = A.shape
    N = B.shape[1]
    absA, absB = np.abs(A).astype(np.float64), np.abs(B).astype(np.float64)
    CC = Cm.copy(); fn(C, dt, "gemm")(RM, NO_TRANS, NO_TRANS, M, N, K, 1.5, A, K, B, N, 0.5, CC, N)
    assert_bound(CC, g["gemm_nn"], 0.5 * np.abs(Cm) + 1.5 * (absA @ absB), K, dt)
    CC = Cm.copy(); fn(C, dt, "symm")(RM, LEFT, LOWER, M, N, 1.5, S, M, Cm.copy(), N, 1.0, CC, N)
    absS = np.abs(S).astype(np.float64)
    assert_bound(CC, g["symm_ll"], np.abs(Cm) + 1.5 * ((absS + absS.T) @ np.abs(Cm)), M, dt)
    CC = np.zeros((M, M), dt); fn(C, dt, "syrk")(RM, UPPER, NO_TRANS, M, K, 1.5, A, K, 1.0, CC, M)
    assert_bound(CC, g["syrk_un"], 1.5 * (absA @ absA.T), K, dt)
