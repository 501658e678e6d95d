"""Writes the golden fixtures under tests/golden/ from the oracle.

The reference ships no golden vectors (its inputs are unseeded, its only
deterministic case is the trsv known-answer matrix).  These fixtures freeze
seeded inputs and the oracle's outputs, strict (same precision, same loop
order as the @proc) and truth (long double), for one small case per kernel
family and precision, so that a later change to the oracle or to the HIP path
is caught against fixed numbers.  Regenerate with

    python tests/golden/make_golden.py

(only after the oracle itself has been re-pinned by tests/test_oracle.py).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import LEFT, LOWER, NO_TRANS, NON_UNIT, RIGHT, TRANS, UNIT, UPPER, Oracle  # noqa: E402

CASES = []


def case(name):
    def deco(f):
        CASES.append((name, f))
        return f
    return deco


def u(rng, shape, dt):
    return rng.uniform(-1, 1, shape).astype(dt)


@case("level1")
def level1(dt, rng):
    n = 1000
    x, y = u(rng, n, dt), u(rng, n, dt)
    o, t = Oracle(dt), Oracle(dt, truth=True)
    out = dict(x=x, y=y)
    out["asum"] = np.array([o.asum(n, x, 1)])
    out["dot"] = np.array([o.dot(n, x, 1, y, 1)])
    out["dot_truth"] = np.array([t.dot(n, x.astype(np.float64), 1, y.astype(np.float64), 1)])
    yy = y.copy(); o.axpy(n, 0.75, x, 1, yy, 1); out["axpy"] = yy
    xx, yy = x.copy(), y.copy(); o.rot(n, xx, 1, yy, 1, 0.6, 0.8); out["rot_x"], out["rot_y"] = xx, yy
    H = np.array([0.3, 1.1, -0.7, 0.45], dt)
    xx, yy = x.copy(), y.copy(); o.rotm(-1, n, xx, 1, yy, 1, H); out["rotm_x"], out["rotm_y"] = xx, yy
    xx = x.copy(); o.scal(n, 2.5, xx, 1); out["scal"] = xx
    return out


@case("gemv")
def gemv(dt, rng):
    m, n = 103, 141
    A, x, y = u(rng, (m, n), dt), u(rng, n, dt), u(rng, m, dt)
    o = Oracle(dt)
    yn = y.copy(); o.gemv(NO_TRANS, m, n, 2.4, 2.4, A, n, x, 1, yn, 1)
    xt = u(rng, m, dt); yt = u(rng, n, dt); ytt = yt.copy()
    o.gemv(TRANS, n, m, 2.4, 2.4, A, n, xt, 1, ytt, 1)
    return dict(A=A, x=x, y=y, gemv_n=yn, xt=xt, yt=yt, gemv_t=ytt)


@case("level2_tri")
def level2_tri(dt, rng):
    n = 100
    A = u(rng, (n, n), dt)
    D = (u(rng, (n, n), dt) / n).astype(dt)
    D[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(dt)
    x, y = u(rng, n, dt), u(rng, n, dt)
    o = Oracle(dt)
    out = dict(A=A, D=D, x=x, y=y)
    yy = y.copy(); o.symv(LOWER, n, 1.5, A, n, x, 1, 0.5, yy, 1); out["symv_l"] = yy
    AA = A.copy(); o.syr2(UPPER, n, 1.2, x, 1, y, 1, AA, n); out["syr2_u"] = AA
    xx = x.copy(); o.trmv(UPPER, TRANS, UNIT, n, xx, 1, A, n); out["trmv_utu"] = xx
    xx = x.copy(); o.trsv(LOWER, NO_TRANS, NON_UNIT, n, xx, 1, D, n); out["trsv_lnn"] = xx
    xx = x.copy(); o.trsv(UPPER, TRANS, NON_UNIT, n, xx, 1, D, n); out["trsv_utn"] = xx
    return out


@case("level3")
def level3(dt, rng):
    M, N, K = 64, 33, 57
    A, B, C = u(rng, (M, K), dt), u(rng, (K, N), dt), u(rng, (M, N), dt)
    o = Oracle(dt)
    out = dict(A=A, B=B, C=C)
    CC = C.copy(); o.mscal(M, N, 0.5, CC, N); o.gemm(NO_TRANS, NO_TRANS, M, N, K, 1.5, A, K, B, N, CC, N)
    out["gemm_nn"] = CC
    S = u(rng, (M, M), dt)
    CC = C.copy(); o.symm(LEFT, LOWER, M, N, 1.5, S, M, C.copy(), N, CC, N); out["S"] = S; out["symm_ll"] = CC
    CC = np.zeros((M, M), dt); o.syrk(UPPER, NO_TRANS, M, K, 1.5, A, K, CC, M); out["syrk_un"] = CC
    return out


def main():
    for name, f in CASES:  # noqa: B007
        for dt in (np.float32, np.float64):
            rng = np.random.default_rng([ord(c) for c in name + np.dtype(dt).name])
            arrays = f(dt, rng)
            path = os.path.join(HERE, f"{name}_{np.dtype(dt).name}.npz")
            np.savez_compressed(path, **arrays)
            print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
