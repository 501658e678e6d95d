"""Level-2 parity: HIP kernels behind the exo_* C-ABI against the oracle.

Sweeps follow the reference's correctness tests:
  gemv  M {77,103,500} x N {3,41,80,101,141,1024} x lda_diff {0,5} x
        incX {1,2} x incY {1,-3}, alpha = beta = 2.4, N/T
        (test/level2/gemv/correctness.cpp:46-77)
  ger   the same shapes plus the skinny M = 2^k, N = 40 sweep (ger/correctness.cpp:60-62)
  symv  N {2,100,321} x alpha {1,2.4} x lda_diff {0,3} x incX {1,3} x
        beta {0,2.4} x incY {1,-2}, U/L (symv/correctness.cpp:37-62)
  syr, syr2  N {2,100,321,1115} x incX {1,2} x lda_diff {0,5}, U/L
  trmv, trsv N {2,100,321} x incX {1,2} x lda_diff {0,5}, U/L x N/T x U/N
and at the bandwidth-set size (8192 x 8192) for gemv, through size-independent
checks.  Results are compared with the strict oracle within the K*eps
bound computed from the float64 magnitudes of the terms.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, NON_UNIT, TRANS, UNIT, UPPER, Oracle
from util import assert_bound, bits_equal, elems, uniform, vec

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
RM = 101


@pytest.fixture(scope="module")
def C():
    from exoblas_amd import cblas, lib
    lib.load()
    return cblas


def fn(C, dt, name):
    return getattr(C, ("exo_s" if dt == np.float32 else "exo_d") + name)


def absvec(buf, n, inc):
    return np.abs(elems(buf, n, inc).astype(np.float64))


GEMV_CASES = list(itertools.product([77, 103, 500], [3, 41, 80, 101, 141, 1024], [0, 5], [1, 2], [1, -3]))


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("M,N,ld_diff,incx,incy", GEMV_CASES)
def test_gemv(C, dt, trans, M, N, ld_diff, incx, incy, rng):
    lda = N + ld_diff
    A = uniform(rng, (M, lda), dt)
    m, n = (M, N) if trans == NO_TRANS else (N, M)  # the wrapper's swap (exo_gemv_wrapper.h:16-17)
    x = vec(rng, n, incx, dt)
    y = vec(rng, m, incy, dt)
    yo = y.copy()
    Oracle(dt).gemv(trans, m, n, 2.4, 2.4, A, lda, x, incx, yo, incy)
    fn(C, dt, "gemv")(RM, trans, M, N, 2.4, A, lda, x, incx, 2.4, y, incy)
    A64 = np.abs(A[:, :N].astype(np.float64))
    opA = A64 if trans == NO_TRANS else A64.T
    mag = yo.astype(np.float64).copy()
    mag[:] = 0
    me = 2.4 * absvec(y, m, incy) + 2.4 * (opA @ absvec(x, n, incx))
    idx = np.arange(m) * abs(incy)
    if incy < 0:
        idx = idx[::-1]
    mag[idx] = me
    assert_bound(y, yo, mag, n, dt, what="gemv")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("M,N,ld_diff,incx,incy", GEMV_CASES[::5] + [(1 << k, 40, 0, 1, 1) for k in range(0, 15, 2)])
def test_ger(C, dt, M, N, ld_diff, incx, incy, rng):
    lda = N + ld_diff
    A = uniform(rng, (M, lda), dt)
    x = vec(rng, M, incx, dt)
    y = vec(rng, N, incy, dt)
    Ao = A.copy()
    Oracle(dt).ger(M, N, 1.7, x, incx, y, incy, Ao, lda)
    fn(C, dt, "ger")(RM, M, N, 1.7, x, incx, y, incy, A, lda)
    assert_bound(A, Ao, np.abs(Ao) + 2, 1, dt, what="ger")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("n,alpha,ld_diff,incx,beta,incy",
                         list(itertools.product([2, 100, 321], [1.0, 2.4], [0, 3], [1, 3], [0.0, 2.4], [1, -2])))
def test_symv(C, dt, uplo, n, alpha, ld_diff, incx, beta, incy, rng):
    lda = n + ld_diff
    A = uniform(rng, (n, lda), dt)
    x = vec(rng, n, incx, dt)
    y = vec(rng, n, incy, dt)
    yo = y.copy()
    Oracle(dt).symv(uplo, n, alpha, A, lda, x, incx, beta, yo, incy)
    fn(C, dt, "symv")(RM, uplo, n, alpha, A, lda, x, incx, beta, y, incy)
    T = np.tril(A[:, :n]) if uplo == LOWER else np.triu(A[:, :n])
    S = np.abs((T + T.T - np.diag(np.diag(T))).astype(np.float64))
    mag = np.zeros(y.shape)
    idx = np.arange(n) * abs(incy)
    if incy < 0:
        idx = idx[::-1]
    mag[idx] = abs(beta) * absvec(yo, n, incy) + abs(beta) * 2 + alpha * (S @ absvec(x, n, incx))
    assert_bound(y, yo, mag, n, dt, what="symv")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("n,incx,ld_diff", list(itertools.product([2, 100, 321, 1115], [1, 2], [0, 5])))
def test_syr_syr2(C, dt, uplo, n, incx, ld_diff, rng):
    lda = n + ld_diff
    A = uniform(rng, (n, lda), dt)
    x = vec(rng, n, incx, dt)
    y = vec(rng, n, incx, dt)
    Ao = A.copy()
    Oracle(dt).syr(uplo, n, 1.2, x, incx, Ao, lda)
    AA = A.copy()
    fn(C, dt, "syr")(RM, uplo, n, 1.2, x, incx, AA, lda)
    assert_bound(AA, Ao, np.abs(Ao) + 2, 1, dt, what="syr")
    mask = np.tril(np.ones((n, n), bool)) if uplo == LOWER else np.triu(np.ones((n, n), bool))
    assert bits_equal(AA[:, :n][~mask], A[:, :n][~mask])
    Ao = A.copy()
    Oracle(dt).syr2(uplo, n, 1.2, x, incx, y, incx, Ao, lda)
    AA = A.copy()
    fn(C, dt, "syr2")(RM, uplo, n, 1.2, x, incx, y, incx, AA, lda)
    assert_bound(AA, Ao, np.abs(Ao) + 4, 2, dt, what="syr2")
    assert bits_equal(AA[:, :n][~mask], A[:, :n][~mask])


TR_CASES = list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS], [UNIT, NON_UNIT], [2, 100, 321], [1, 2], [0, 5]))


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag,n,incx,ld_diff", TR_CASES)
def test_trmv(C, dt, uplo, trans, diag, n, incx, ld_diff, rng):
    lda = n + ld_diff
    A = uniform(rng, (n, lda), dt)
    x = vec(rng, n, incx, dt)
    xo = x.copy()
    Oracle(dt).trmv(uplo, trans, diag, n, xo, incx, A, lda)
    fn(C, dt, "trmv")(RM, uplo, trans, diag, n, A, lda, x, incx)
    T = np.abs(np.tril(A[:, :n]) if uplo == LOWER else np.triu(A[:, :n])).astype(np.float64)
    if trans != NO_TRANS:
        T = T.T
    mag = np.zeros(x.shape)
    idx = np.arange(n) * incx
    mag[idx] = T @ np.abs(x[idx].astype(np.float64)) + 2 * np.abs(x[idx])
    assert_bound(x, xo, mag, n, dt, what="trmv")


@pytest.mark.parametrize("uplo,diag", [(UPPER, UNIT), (LOWER, NON_UNIT)])
def test_trmv_tall(C, uplo, diag, rng):
    """n = 8200: the one-launch, two-rows-per-wave row path of the triangular mat-vec."""
    test_trmv(C, np.float32, uplo, NO_TRANS, diag, 8200, 1, 3, rng)


def dd_matrix(rng, n, lda, dt):
    """Diagonally dominant: off-diagonal U[-1,1]/n, diagonal n + U[0,1]."""
    A = (rng.uniform(-1, 1, (n, lda)) / n).astype(dt)
    A[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(dt)
    return A


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag,n,incx,ld_diff", TR_CASES)
def test_trsv(C, dt, uplo, trans, diag, n, incx, ld_diff, rng):
    lda = n + ld_diff
    A = dd_matrix(rng, n, lda, dt) if diag == NON_UNIT else (uniform(rng, (n, lda), dt) / n).astype(dt)
    x = vec(rng, n, incx, dt)
    xo = x.copy()
    Oracle(dt).trsv(uplo, trans, diag, n, xo, incx, A, lda)
    fn(C, dt, "trsv")(RM, uplo, trans, diag, n, A, lda, x, incx)
    assert_bound(x, xo, np.abs(xo) + 1, n, dt, what="trsv")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS],
                                                                  [UNIT, NON_UNIT])))
@pytest.mark.parametrize("n,incx,ld_diff", [(1024, 1, 0), (1500, 1, 3), (2600, 2, 5)])
def test_trsv_long(C, dt, uplo, trans, diag, n, incx, ld_diff, rng):
    """Long solves: chains of 16 to 41 dependent 64-row blocks."""
    test_trsv(C, dt, uplo, trans, diag, n, incx, ld_diff, rng)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [2, 257, 1000, 1100, 2049])
def test_trsv_known_answer(C, dt, n):
    # test/level2/trsv/correctness.cpp:20-25: every A[i][j] = 2.0, x[i] = i
    A = np.full((n, n), 2.0, dt)
    x = np.arange(n, dtype=dt)
    fn(C, dt, "trsv")(RM, LOWER, NO_TRANS, NON_UNIT, n, A, n, x, 1)
    assert x[0] == 0 and np.all(x[1:] == 0.5)
    x = np.arange(n, dtype=dt)
    fn(C, dt, "trsv")(RM, LOWER, NO_TRANS, UNIT, n, A, n, x, 1)
    assert np.all(x == np.arange(n) % 2)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
def test_gemv_bandwidth_size(C, dt, trans, rng):
    """8192 x 8192, lda 8192: against float64 numpy on the device path."""
    import torch
    n = 8192
    A = uniform(rng, (n, n), dt)
    x = uniform(rng, n, dt)
    y = uniform(rng, n, dt)
    dA, dx, dy = (torch.from_numpy(a).cuda() for a in (A, x, y))
    fn(C, dt, "gemv")(RM, trans, n, n, 0.5, dA, n, dx, 1, 1.5, dy, 1)
    A64 = A.astype(np.float64)
    opA = A64 if trans == NO_TRANS else A64.T
    want = 1.5 * y + 0.5 * (opA @ x.astype(np.float64))
    mag = 1.5 * np.abs(y) + 0.5 * (np.abs(opA) @ np.abs(x.astype(np.float64)))
    assert_bound(dy.cpu().numpy(), want, mag, n, dt, c=1.0, what="gemv 8192")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("M,N,lda", [(4099, 3001, 3004), (4100, 5, 8), (5000, 4097, 4100), (4096, 8191, 8192)])
def test_gemv_n_long_rows(C, dt, M, N, lda, rng):
    """M >= 4096 rows with a contiguous x: whole-workgroup rows with x held in
    registers, an odd row count and a partial last vector of columns."""
    import torch
    A = uniform(rng, (M, lda), dt)
    x = uniform(rng, N, dt)
    y = uniform(rng, M, dt)
    dA, dx, dy = (torch.from_numpy(a).cuda() for a in (A, x, y))
    fn(C, dt, "gemv")(RM, NO_TRANS, M, N, -0.75, dA, lda, dx, 1, 1.25, dy, 1)
    A64 = A[:, :N].astype(np.float64)
    want = 1.25 * y + -0.75 * (A64 @ x.astype(np.float64))
    mag = 1.25 * np.abs(y) + 0.75 * (np.abs(A64) @ np.abs(x.astype(np.float64)))
    assert_bound(dy.cpu().numpy(), want, mag, N, dt, c=1.0, what="gemv long rows")
