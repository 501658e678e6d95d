"""Pins the CPU oracle (oracle/exo_oracle.cpp) before anything is checked
against it.

The reference ships no golden vectors; its only deterministic input is the
trsv known-answer matrix (test/level2/trsv/correctness.cpp:20-25).  Each
restated routine is therefore held to something independent of its own
formula:
  * closed forms (copy/swap bit-exact, trsv on the all-2 matrix, sums of
    integers that fp32/fp64 represent exactly, the special-case early returns),
  * library routines (scipy.linalg.blas / numpy float64 products) on seeded
    inputs, within the K*eps forward-error bound,
  * brute-force float64 loops on tiny inputs written here from the
    mathematical definition (not from the spec's loop order),
and each check is shown to reject a planted mistake (wrong sign, transposed
operand, other triangle) so it can fail.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.linalg.blas as sb

from oracle import (LEFT, LOWER, NO_TRANS, NON_UNIT, RIGHT, TRANS, UNIT, UPPER, Oracle)
from util import assert_bound, bits_equal, elems, eps, uniform, vec

DT = [np.float32, np.float64]


def O(dt):
    return Oracle(dt)


# ---------------------------------------------------------------- level 1

@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,inc", [(0, 1), (2, 1), (321, 1), (1000, 3), (321, -2)])
def test_copy_swap_bit_exact(dt, n, inc, rng):
    x = vec(rng, n, inc, dt)
    y = vec(rng, n, inc, dt)
    y0 = y.copy()
    O(dt).copy(n, x, inc, y, inc)
    assert bits_equal(elems(y, n, inc), elems(x, n, inc))
    if n:
        mask = np.ones(y.size, bool)
        mask[np.arange(n) * abs(inc)] = False
        assert bits_equal(y[mask], y0[mask])  # padding untouched
    x = vec(rng, n, inc, dt)
    y = vec(rng, n, inc, dt)
    x0, y0 = x.copy(), y.copy()
    O(dt).swap(n, x, inc, y, inc)
    assert bits_equal(elems(x, n, inc), elems(y0, n, inc))
    assert bits_equal(elems(y, n, inc), elems(x0, n, inc))


@pytest.mark.parametrize("dt", DT)
def test_asum_dot_integer_closed_form(dt):
    # sums of small integers are exact in fp32 below 2^24
    n = 1000
    x = (np.arange(n) % 7 - 3).astype(dt)
    y = (np.arange(n) % 5 - 2).astype(dt)
    o = O(dt)
    assert o.asum(n, x, 1) == float(np.sum(np.abs(x.astype(np.int64))))
    assert o.dot(n, x, 1, y, 1) == float(np.sum(x.astype(np.int64) * y.astype(np.int64)))
    assert o.asum(0, x, 1) == 0.0


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,inc", [(2, 1), (321, 1), (1000, 1), (1000, 3)])
def test_level1_vs_scipy(dt, n, inc, rng):
    p = "s" if dt == np.float32 else "d"
    o = O(dt)
    x = vec(rng, n, inc, dt)
    y = vec(rng, n, inc, dt)
    xe, ye = elems(x, n, inc).astype(np.float64), elems(y, n, inc).astype(np.float64)
    # asum / dot against the library and the K*eps bound around float64
    assert_bound(o.asum(n, x, inc), getattr(sb, p + "asum")(x, n=n, incx=inc), np.sum(np.abs(xe)), n, dt,
                 what="asum")
    assert_bound(o.dot(n, x, inc, y, inc), np.dot(xe, ye), np.sum(np.abs(xe * ye)), n, dt, what="dot")
    assert_bound(o.dot(n, x, inc, y, inc), getattr(sb, p + "dot")(x, y, n=n, incx=inc, incy=inc),
                 np.sum(np.abs(xe * ye)), n, dt, what="dot/scipy")
    # axpy
    a = 0.75
    yy = y.copy()
    o.axpy(n, a, x, inc, yy, inc)
    assert_bound(elems(yy, n, inc), ye + a * xe, np.abs(ye) + np.abs(a * xe), 1, dt, what="axpy")
    yy = y.copy()
    o.axpy_alpha_1(n, x, inc, yy, inc)
    assert_bound(elems(yy, n, inc), ye + xe, np.abs(ye) + np.abs(xe), 1, dt, what="axpy1")
    # scal
    xx = x.copy()
    o.scal(n, 2.5, xx, inc)
    assert_bound(elems(xx, n, inc), 2.5 * xe, np.abs(2.5 * xe), 1, dt, what="scal")
    xx = x.copy()
    o.scal_alpha_0(n, xx, inc)
    assert np.all(elems(xx, n, inc) == 0)
    # rot, c = 0.6, s = 0.8
    c, s = 0.6, 0.8
    xx, yy = x.copy(), y.copy()
    o.rot(n, xx, inc, yy, inc, c, s)
    ref_x, ref_y = c * xe + s * ye, -s * xe + c * ye
    mag = np.abs(c * xe) + np.abs(s * ye)
    assert_bound(elems(xx, n, inc), ref_x, mag, 2, dt, what="rot x")
    assert_bound(elems(yy, n, inc), ref_y, mag, 2, dt, what="rot y")
    xs, ys = getattr(sb, p + "rot")(x.copy(), y.copy(), c, s, n=n, incx=inc, incy=inc)
    assert_bound(elems(xx, n, inc), elems(xs, n, inc), mag, 2, dt, what="rot/scipy x")
    assert_bound(elems(yy, n, inc), elems(ys, n, inc), mag, 2, dt, what="rot/scipy y")
    # planted mistake: the other rotation direction must fail the same check
    xx, yy = x.copy(), y.copy()
    o.rot(n, xx, inc, yy, inc, c, -s)
    with pytest.raises(AssertionError):
        assert_bound(elems(yy, n, inc), ref_y, mag, 2, dt, what="planted")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("flag", [-1.0, 0.0, 1.0])
def test_rotm_vs_scipy(dt, flag, rng):
    # netlib param P = [flag, h11, h21, h12, h22]; the reference wrapper passes
    # H = {P1, P3, P2, P4} = {h11, h12, h21, h22} (exo_rotm_wrapper.h:8)
    p = "s" if dt == np.float32 else "d"
    n = 321
    x = uniform(rng, n, dt)
    y = uniform(rng, n, dt)
    P = np.array([flag, 0.3, -0.7, 1.1, 0.45], dtype=dt)
    H = np.array([P[1], P[3], P[2], P[4]], dtype=dt)
    xx, yy = x.copy(), y.copy()
    O(dt).rotm(int(flag), n, xx, 1, yy, 1, H)
    xs, ys = getattr(sb, p + "rotm")(x.copy(), y.copy(), P)
    mag = np.abs(x) * 1.2 + np.abs(y) * 1.2
    assert_bound(xx, xs, mag, 2, dt, what="rotm x")
    assert_bound(yy, ys, mag, 2, dt, what="rotm y")


def test_dsdot_sdsdot(rng):
    n = 1000
    x = uniform(rng, n, np.float32)
    y = uniform(rng, n, np.float32)
    o = O(np.float32)
    exact = math.fsum((x.astype(np.float64) * y.astype(np.float64)).tolist())  # fp32*fp32 exact in fp64
    mag = float(np.sum(np.abs(x.astype(np.float64) * y)))
    assert abs(o.dsdot(n, x, 1, y, 1) - exact) <= n * 2.0 ** -53 * mag
    assert abs(o.dsdot(n, x, 1, y, 1, truth=True) - exact) <= 4 * 2.0 ** -53 * mag
    sbv = np.float32(0.125)
    assert o.sdsdot(n, float(sbv), x, 1, y, 1) == np.float32(np.float64(exact) + 0.125) or \
        abs(o.sdsdot(n, float(sbv), x, 1, y, 1) - (exact + 0.125)) <= 2 ** -23 * (abs(exact) + 1)
    # integer closed form: exact in fp32 after the fp64 sum
    xi = (np.arange(n) % 9 - 4).astype(np.float32)
    yi = (np.arange(n) % 3 - 1).astype(np.float32)
    assert o.dsdot(n, xi, 1, yi, 1) == float(np.sum(xi.astype(np.int64) * yi.astype(np.int64)))


# ---------------------------------------------------------------- level 2

def _tri(A, uplo, unit=False):
    T = np.tril(A) if uplo == LOWER else np.triu(A)
    if unit:
        np.fill_diagonal(T, 1.0)
    return T


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("m,n,lda_diff", [(77, 3, 0), (103, 141, 5), (500, 80, 0)])
def test_gemv(dt, trans, m, n, lda_diff, rng):
    # callee view: A is m x n (NoTrans) or n x m (Trans), y has m elements
    rows, cols = (m, n) if trans == NO_TRANS else (n, m)
    lda = cols + lda_diff
    A = uniform(rng, (rows, lda), dt)
    x = uniform(rng, n, dt)
    y = uniform(rng, m, dt)
    alpha = beta = 2.4
    yy = y.copy()
    O(dt).gemv(trans, m, n, alpha, beta, A, lda, x, 1, yy, 1)
    A64 = A[:, :cols].astype(np.float64)
    opA = A64 if trans == NO_TRANS else A64.T
    ref = beta * y.astype(np.float64) + alpha * (opA @ x.astype(np.float64))
    mag = np.abs(beta * y) + np.abs(alpha) * (np.abs(opA) @ np.abs(x.astype(np.float64)))
    assert_bound(yy, ref, mag, n, dt, what="gemv")
    # library routine (column-major view of the row-major matrix flips trans)
    p = "s" if dt == np.float32 else "d"
    ys = getattr(sb, p + "gemv")(alpha, A[:, :cols].T, x, beta, y.copy(), trans=0 if trans == TRANS else 1)
    assert_bound(yy, ys, mag, n, dt, what="gemv/scipy")
    # planted mistake: the transposed operand fails when the shapes allow it
    if m == n or True:
        sq = min(m, n)
        if sq > 1:
            B = uniform(rng, (sq, sq), dt)
            y1 = np.zeros(sq, dt)
            O(dt).gemv(NO_TRANS, sq, sq, 1.0, 0.0, B, sq, x[:sq].copy(), 1, y1, 1)
            wrong = B.astype(np.float64).T @ x[:sq].astype(np.float64)
            with pytest.raises(AssertionError):
                assert_bound(y1, wrong, np.abs(B.astype(np.float64)).T @ np.abs(x[:sq]), sq, dt, c=1.0)


def test_oracle_rejects_short_buffers():
    # CBLAS sizes passed where the callee's (swapped for Trans) are expected:
    # y would receive 30 values in a buffer of 20.  The binding refuses
    # before the C restatement writes past the array.
    o = O(np.float32)
    A = np.zeros((30, 20), np.float32)
    x = np.zeros(30, np.float32)
    y = np.zeros(20, np.float32)
    with pytest.raises(ValueError):
        o.gemv(TRANS, 30, 20, 1.0, 0.0, A, 20, x, 1, y, 1)
    o.gemv(TRANS, 20, 30, 1.0, 0.0, A, 20, x, 1, y, 1)  # the callee view fits
    with pytest.raises(ValueError):
        o.axpy(5, 1.0, x, 1, y[:4].copy(), -1)
    with pytest.raises(ValueError):
        o.gemm(NO_TRANS, NO_TRANS, 30, 20, 21, 1.0, A, 20, A, 20, A, 20)


@pytest.mark.parametrize("dt", DT)
def test_ger(dt, rng):
    m, n, lda = 103, 41, 46
    A = uniform(rng, (m, lda), dt)
    x = uniform(rng, m, dt)
    y = uniform(rng, n, dt)
    AA = A.copy()
    O(dt).ger(m, n, 1.5, x, 1, y, 1, AA, lda)
    ref = A.astype(np.float64).copy()
    ref[:, :n] += 1.5 * np.outer(x.astype(np.float64), y)
    assert_bound(AA, ref, np.abs(ref) + 1.5, 2, dt, what="ger")
    assert bits_equal(AA[:, n:], A[:, n:])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
def test_symv_syr_syr2(dt, uplo, rng):
    n, lda = 141, 146
    A = uniform(rng, (n, lda), dt)
    x = uniform(rng, n, dt)
    y = uniform(rng, n, dt)
    o = O(dt)
    # symv: only the `uplo` triangle is read
    T = _tri(A[:, :n].astype(np.float64), uplo)
    S = T + T.T - np.diag(np.diag(T))
    yy = y.copy()
    o.symv(uplo, n, 2.4, A, lda, x, 1, 2.4, yy, 1)
    ref = 2.4 * y + 2.4 * (S @ x.astype(np.float64))
    mag = 2.4 * np.abs(y) + 2.4 * (np.abs(S) @ np.abs(x.astype(np.float64)))
    assert_bound(yy, ref, mag, n, dt, what="symv")
    # planted mistake: the other triangle
    T2 = _tri(A[:, :n].astype(np.float64), UPPER if uplo == LOWER else LOWER)
    S2 = T2 + T2.T - np.diag(np.diag(T2))
    with pytest.raises(AssertionError):
        assert_bound(yy, 2.4 * y + 2.4 * (S2 @ x.astype(np.float64)), mag, n, dt)
    # syr: only the triangle is written
    AA = A.copy()
    o.syr(uplo, n, 1.3, x, 1, AA, lda)
    upd = 1.3 * np.outer(x.astype(np.float64), x)
    mask = np.tril(np.ones((n, n), bool)) if uplo == LOWER else np.triu(np.ones((n, n), bool))
    ref = A[:, :n].astype(np.float64) + np.where(mask, upd, 0.0)
    assert_bound(AA[:, :n], ref, np.abs(ref) + 2, 2, dt, what="syr")
    assert bits_equal(AA[:, :n][~mask], A[:, :n][~mask])
    # syr2
    AA = A.copy()
    o.syr2(uplo, n, 1.3, x, 1, y, 1, AA, lda)
    upd = 1.3 * (np.outer(x.astype(np.float64), y) + np.outer(y.astype(np.float64), x))
    ref = A[:, :n].astype(np.float64) + np.where(mask, upd, 0.0)
    assert_bound(AA[:, :n], ref, np.abs(ref) + 3, 3, dt, what="syr2")
    assert bits_equal(AA[:, :n][~mask], A[:, :n][~mask])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("diag", [NON_UNIT, UNIT])
def test_trmv_trsv(dt, uplo, trans, diag, rng):
    n, lda = 101, 104
    A = uniform(rng, (n, lda), dt)
    x = uniform(rng, n, dt)
    o = O(dt)
    T = _tri(A[:, :n].astype(np.float64), uplo, unit=diag == UNIT)
    opT = T if trans == NO_TRANS else T.T
    xx = x.copy()
    o.trmv(uplo, trans, diag, n, xx, 1, A, lda)
    mag = np.abs(opT) @ np.abs(x.astype(np.float64)) + np.abs(x)
    assert_bound(xx, opT @ x.astype(np.float64), mag, n, dt, what="trmv")
    # trsv on a well-conditioned matrix (diagonal n + U[0,1], off-diagonal U[-1,1]/n)
    B = (uniform(rng, (n, lda), dt) / n).astype(dt)
    B[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(dt)
    Tb = _tri(B[:, :n].astype(np.float64), uplo, unit=diag == UNIT)
    opB = Tb if trans == NO_TRANS else Tb.T
    b = uniform(rng, n, dt)
    xx = b.copy()
    o.trsv(uplo, trans, diag, n, xx, 1, B, lda)
    ref = np.linalg.solve(opB, b.astype(np.float64))
    assert_bound(xx, ref, np.abs(ref) + 1, n, dt, what="trsv")
    p = "s" if dt == np.float32 else "d"
    # scipy: column-major view of the row-major matrix is its transpose
    xs = getattr(sb, p + "trsv")(B[:, :n].T, b.copy(), lower=int(uplo == UPPER),
                                 trans=0 if trans == TRANS else 1, diag=int(diag == UNIT))
    assert_bound(xx, xs, np.abs(ref) + 1, n, dt, what="trsv/scipy")


@pytest.mark.parametrize("dt", DT)
def test_trsv_known_answer(dt):
    # test/level2/trsv/correctness.cpp:20-25: A[i][j] = 2.0 + j/N (integer
    # division, so every element is 2.0), x[i] = i.
    n = 257
    A = np.full((n, n), 2.0, dt)
    o = O(dt)
    x = np.arange(n, dtype=dt)
    o.trsv(LOWER, NO_TRANS, NON_UNIT, n, x, 1, A, n)
    assert x[0] == 0 and np.all(x[1:] == 0.5)
    x = np.arange(n, dtype=dt)
    o.trsv(LOWER, NO_TRANS, UNIT, n, x, 1, A, n)
    assert np.all(x == (np.arange(n) % 2))
    # upper / transposed forms are the mirrored closed forms
    x = np.arange(n, dtype=dt)[::-1].copy()
    o.trsv(UPPER, NO_TRANS, NON_UNIT, n, x, 1, A, n)
    assert x[-1] == 0 and np.all(x[:-1] == 0.5)
    x = np.arange(n, dtype=dt)
    o.trsv(UPPER, TRANS, NON_UNIT, n, x, 1, A, n)
    assert x[0] == 0 and np.all(x[1:] == 0.5)


# ---------------------------------------------------------------- level 3

@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("ta", [NO_TRANS, TRANS])
@pytest.mark.parametrize("tb", [NO_TRANS, TRANS])
@pytest.mark.parametrize("M,N,K", [(1, 7, 32), (7, 32, 1), (64, 33, 57)])
def test_gemm(dt, ta, tb, M, N, K, rng):
    A = uniform(rng, (M, K) if ta == NO_TRANS else (K, M), dt)
    B = uniform(rng, (K, N) if tb == NO_TRANS else (N, K), dt)
    C = uniform(rng, (M, N), dt)
    o = O(dt)
    CC = C.copy()
    o.mscal(M, N, 18.0, CC, N)
    assert bits_equal(CC, (C * dt(18.0)).astype(dt))
    o.gemm(ta, tb, M, N, K, 13.0, A, A.shape[1], B, B.shape[1], CC, N)
    opA = A.astype(np.float64) if ta == NO_TRANS else A.T.astype(np.float64)
    opB = B.astype(np.float64) if tb == NO_TRANS else B.T.astype(np.float64)
    ref = 18.0 * C + 13.0 * (opA @ opB)
    mag = 18.0 * np.abs(C) + 13.0 * (np.abs(opA) @ np.abs(opB))
    assert_bound(CC, ref, mag, K, dt, what="gemm")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
def test_symm(dt, side, uplo, rng):
    M, N = 33, 20
    na = M if side == LEFT else N
    A = uniform(rng, (na, na + 5), dt)
    B = uniform(rng, (M, N), dt)
    C = np.zeros((M, N), dt)
    O(dt).symm(side, uplo, M, N, 1.5, A, na + 5, B, N, C, N)
    T = _tri(A[:, :na].astype(np.float64), uplo)
    S = T + T.T - np.diag(np.diag(T))
    ref = 1.5 * (S @ B if side == LEFT else B @ S)
    mag = 1.5 * (np.abs(S) @ np.abs(B) if side == LEFT else np.abs(B) @ np.abs(S))
    assert_bound(C, ref, mag, na, dt, what="symm")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
def test_syrk_trmscal(dt, uplo, trans, rng):
    N, K = 37, 19
    A = uniform(rng, (N, K) if trans == NO_TRANS else (K, N), dt)
    C = uniform(rng, (N, N), dt)
    o = O(dt)
    CC = C.copy()
    o.trmscal(uplo, N, 18.0, CC, N)
    mask = np.tril(np.ones((N, N), bool)) if uplo == LOWER else np.triu(np.ones((N, N), bool))
    assert bits_equal(CC[mask], (C[mask] * dt(18.0)).astype(dt))
    assert bits_equal(CC[~mask], C[~mask])
    o.syrk(uplo, trans, N, K, 13.0, A, A.shape[1], CC, N)
    A64 = A.astype(np.float64)
    G = A64 @ A64.T if trans == NO_TRANS else A64.T @ A64
    Ga = np.abs(A64) @ np.abs(A64).T if trans == NO_TRANS else np.abs(A64).T @ np.abs(A64)
    ref = np.where(mask, 18.0 * C + 13.0 * G, C)
    assert_bound(CC, ref, 18.0 * np.abs(C) + 13.0 * Ga, K, dt, what="syrk")
    assert bits_equal(CC[~mask], C[~mask])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
def test_syr2k(dt, uplo, trans, rng):
    N, K = 33, 21
    shape = (N, K) if trans == NO_TRANS else (K, N)
    A, B = uniform(rng, shape, dt), uniform(rng, shape, dt)
    C = uniform(rng, (N, N), dt)
    CC = C.copy()
    O(dt).syr2k(uplo, trans, N, K, 13.0, A, A.shape[1], B, B.shape[1], CC, N)
    mask = np.tril(np.ones((N, N), bool)) if uplo == LOWER else np.triu(np.ones((N, N), bool))
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    G = A64 @ B64.T + B64 @ A64.T if trans == NO_TRANS else A64.T @ B64 + B64.T @ A64
    aA, aB = np.abs(A64), np.abs(B64)
    Ga = aA @ aB.T + aB @ aA.T if trans == NO_TRANS else aA.T @ aB + aB.T @ aA
    ref = np.where(mask, C + 13.0 * G, C)
    assert_bound(CC, ref, np.abs(C) + 13.0 * Ga, 2 * K, dt, what="syr2k")
    assert bits_equal(CC[~mask], C[~mask])
    # library routine: scipy's syr2k on the same (logical) operands
    p = "s" if dt == np.float32 else "d"
    lib = getattr(sb, p + "syr2k")(13.0, A, B, beta=1.0, c=C.copy(), trans=int(trans != NO_TRANS),
                                   lower=int(uplo == LOWER))
    assert_bound(CC[mask], lib[mask], (np.abs(C) + 13.0 * Ga)[mask], 2 * K, dt, what="syr2k/scipy")
    # planted mistake: one product only (A B^T without B A^T) fails the bound
    G1 = A64 @ B64.T if trans == NO_TRANS else A64.T @ B64
    with pytest.raises(AssertionError):
        assert_bound(CC, np.where(mask, C + 13.0 * (G1 + G1), C), np.abs(C) + 13.0 * Ga, 2 * K, dt, c=1.0)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("diag", [NON_UNIT, UNIT])
@pytest.mark.parametrize("m,n,ld_diff", [(4, 8, 0), (29, 15, 5)])
def test_trmm(dt, uplo, trans, diag, m, n, ld_diff, rng):
    lda, ldb = m + ld_diff, n + ld_diff
    A = uniform(rng, (m, lda), dt)
    B = uniform(rng, (m, ldb), dt)
    BB = B.copy()
    O(dt).trmm(uplo, trans, diag, m, n, 1.4, A, lda, BB, ldb)
    T = np.tril(A[:, :m].astype(np.float64)) if uplo == LOWER else np.triu(A[:, :m].astype(np.float64))
    if diag == UNIT:
        np.fill_diagonal(T, 1.0)
    opT = T if trans == NO_TRANS else T.T
    B64 = B[:, :n].astype(np.float64)
    ref = 1.4 * (opT @ B64)
    mag = 1.4 * (np.abs(opT) @ np.abs(B64))
    assert_bound(BB[:, :n], ref, mag, m, dt, what="trmm")
    assert bits_equal(BB[:, n:], B[:, n:])  # padding columns untouched
    # library routine: scipy's trmm (logical operands, Left side)
    p = "s" if dt == np.float32 else "d"
    lib = getattr(sb, p + "trmm")(1.4, np.ascontiguousarray(A[:, :m]), np.ascontiguousarray(B[:, :n]), side=0,
                                  lower=int(uplo == LOWER), trans_a=int(trans != NO_TRANS), diag=int(diag == UNIT))
    assert_bound(BB[:, :n], lib, mag, m, dt, what="trmm/scipy")
    # planted mistake: the other triangle fails the bound
    Tw = np.triu(A[:, :m].astype(np.float64)) if uplo == LOWER else np.tril(A[:, :m].astype(np.float64))
    if diag == UNIT:
        np.fill_diagonal(Tw, 1.0)
    wrong = 1.4 * ((Tw if trans == NO_TRANS else Tw.T) @ B64)
    if m > 1:
        with pytest.raises(AssertionError):
            assert_bound(BB[:, :n], wrong, mag, m, dt, c=1.0)


# ---------------------------------------------------------------- banded level 2
# Dense matrices from the row-major band layout (CBLAS definition, written here
# independently of the spec's loops): general A(i, j) = a[i, kl + j - i];
# upper-stored A(i, j>=i) = a[i, j - i]; lower-stored A(i, j<=i) = a[i, k + j - i].

def _band_dense(a, m, n, lo, hi, c0):
    D = np.zeros((m, n))
    for i in range(m):
        for j in range(max(0, i - lo), min(n, i + hi + 1)):
            D[i, j] = a[i, c0 + j - i]
    return D


def _fortran_band(D, kl, ku):
    """Column-major band array of netlib BLAS: ab[ku + i - j, j] = D(i, j)."""
    m, n = D.shape
    ab = np.zeros((kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(m, j + kl + 1)):
            ab[ku + i - j, j] = D[i, j]
    return ab


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("m,n,kl,ku", [(2, 2, 1, 1), (15, 20, 0, 5), (40, 33, 7, 3), (29, 29, 28, 28)])
def test_gbmv(dt, m, n, kl, ku, rng):
    lda = kl + ku + 1
    a = uniform(rng, (m, lda), dt)
    x, y = uniform(rng, n, dt), uniform(rng, m, dt)
    yy = y.copy()
    O(dt).gbmv(m, n, kl, ku, 1.2, 1.2, a, lda, x, 1, yy, 1)
    D = _band_dense(a.astype(np.float64), m, n, kl, ku, kl)
    ref = 1.2 * y + 1.2 * (D @ x)
    mag = 1.2 * np.abs(y) + 1.2 * (np.abs(D) @ np.abs(x.astype(np.float64)))
    assert_bound(yy, ref, mag, kl + ku + 1, dt, what="gbmv")
    p = "s" if dt == np.float32 else "d"
    if m >= kl + ku + 1:  # scipy's binding requires it
        lib = getattr(sb, p + "gbmv")(m, n, kl, ku, 1.2, _fortran_band(D, kl, ku).astype(dt), x, beta=1.2,
                                      y=y.copy())
        assert_bound(yy, lib, mag, kl + ku + 1, dt, what="gbmv/scipy")
    with pytest.raises(AssertionError):  # planted: the band read one column off
        Dw = _band_dense(np.roll(a.astype(np.float64), 1, axis=1), m, n, kl, ku, kl)
        assert_bound(yy, 1.2 * y + 1.2 * (Dw @ x), mag, kl + ku + 1, dt, c=1.0)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("n,k,inc", [(2, 1, 1), (100, 0, 1), (100, 7, -2), (33, 32, 3)])
def test_sbmv(dt, uplo, n, k, inc, rng):
    A = uniform(rng, (n, k + 1 + 3), dt)
    x, y = vec(rng, n, inc, dt), vec(rng, n, inc, dt)
    yy = y.copy()
    o = O(dt)
    o.sbmv_scal_y(n, 0.5, yy, inc)
    o.sbmv(uplo, n, k, 2.4, A, A.shape[1], x, inc, yy, inc)
    A64 = A.astype(np.float64)
    T = _band_dense(A64, n, n, 0, k, 0) if uplo == UPPER else _band_dense(A64, n, n, k, 0, k)
    S = T + T.T - np.diag(np.diag(T))
    xe, ye = elems(x, n, inc).astype(np.float64), elems(y, n, inc).astype(np.float64)
    ref = 0.5 * ye + 2.4 * (S @ xe)
    mag = 0.5 * np.abs(ye) + 2.4 * (np.abs(S) @ np.abs(xe))
    assert_bound(elems(yy, n, inc), ref, mag, 2 * k + 2, dt, what="sbmv")
    p = "s" if dt == np.float32 else "d"
    ab = _fortran_band(T, 0, k)[:, :] if uplo == UPPER else _fortran_band(T, k, 0)
    lib = getattr(sb, p + "sbmv")(k, 2.4, ab.astype(dt), xe.astype(dt), beta=0.5, y=ye.astype(dt),
                                  lower=int(uplo == LOWER))
    assert_bound(elems(yy, n, inc), lib, mag, 2 * k + 2, dt, what="sbmv/scipy")
    if k > 0:
        with pytest.raises(AssertionError):  # planted: the stored triangle only (no mirror)
            assert_bound(elems(yy, n, inc), 0.5 * ye + 2.4 * (T @ xe), mag, 2 * k + 2, dt, c=1.0)


def _tri_band(rng, n, k, uplo, dt, dominant):
    A = uniform(rng, (n, k + 1 + 2), dt)
    if dominant:
        A[:, :k + 1] /= (k + 1)
        dcol = 0 if uplo == UPPER else k
        A[:, dcol] = (1.0 + rng.uniform(0, 1, n)).astype(dt)
    A64 = A.astype(np.float64)
    T = _band_dense(A64, n, n, 0, k, 0) if uplo == UPPER else _band_dense(A64, n, n, k, 0, k)
    return A, T


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("diag", [NON_UNIT, UNIT])
@pytest.mark.parametrize("n,k,inc", [(2, 1, 1), (100, 6, 3), (57, 56, -1)])
def test_tbmv_tbsv(dt, uplo, trans, diag, n, k, inc, rng):
    p = "s" if dt == np.float32 else "d"
    o = O(dt)
    A, T = _tri_band(rng, n, k, uplo, dt, dominant=True)
    Tu = T.copy()
    if diag == UNIT:
        np.fill_diagonal(Tu, 1.0)
    opT = Tu if trans == NO_TRANS else Tu.T
    lower = int(uplo == LOWER)
    ab = (_fortran_band(T, 0, k) if uplo == UPPER else _fortran_band(T, k, 0)).astype(dt)
    tr = int(trans != NO_TRANS)
    # tbmv
    x = vec(rng, n, inc, dt)
    xx = x.copy()
    o.tbmv(uplo, trans, diag, n, k, xx, inc, A, A.shape[1])
    xe = elems(x, n, inc).astype(np.float64)
    mag = np.abs(opT) @ np.abs(xe)
    assert_bound(elems(xx, n, inc), opT @ xe, mag, k + 1, dt, what="tbmv")
    lib = getattr(sb, p + "tbmv")(k, ab, xe.astype(dt), lower=lower, trans=tr, diag=int(diag == UNIT))
    assert_bound(elems(xx, n, inc), lib, mag, k + 1, dt, what="tbmv/scipy")
    if k > 0:
        with pytest.raises(AssertionError):  # planted: the transposed band
            assert_bound(elems(xx, n, inc), opT.T @ xe, mag, k + 1, dt, c=1.0)
    # tbsv: solve, then the residual of the float64 system is small
    b = vec(rng, n, inc, dt)
    bb = b.copy()
    o.tbsv(uplo, trans, diag, n, k, bb, inc, A, A.shape[1])
    be = elems(b, n, inc).astype(np.float64)
    want = np.linalg.solve(opT, be)
    got = elems(bb, n, inc).astype(np.float64)
    assert np.max(np.abs(got - want)) <= 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1), "tbsv"
    lib = getattr(sb, p + "tbsv")(k, ab, be.astype(dt), lower=lower, trans=tr, diag=int(diag == UNIT))
    assert np.max(np.abs(got - lib)) <= 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1), "tbsv/scipy"
    if k > 0:  # planted: solving with the transposed band misses
        wrong = np.linalg.solve(opT.T, be)
        assert np.max(np.abs(got - wrong)) > 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1)
