"""Parity of the banded level-2 kernels (the reference's exo_gbmv, exo_sbmv,
exo_tbmv and exo_tbsv library targets) behind the exo_* C-ABI, against the
oracle, through exoblas_amd.cblas (which mirrors the wrappers, including
sbmv's separate beta scaling and tbsv's stride_any-only Trans forms).

Sweeps follow the reference's correctness tests: gbmv sizes (2, 2, 1, 1),
(15, 20, 0, 5), (256, 256, 255, 255) with alpha = beta = 1.2
(test/level2/sgbmv/correctness.cpp:36-46), plus strided and negative
increments; sbmv N in {2, 100, 321}, K in {0, 1, 50, N-1}, lda + {0, 3},
incX in {1, 3}, beta in {0, 1}, incY in {1, -2} (ssbmv/correctness.cpp:39-70);
tbmv N in {2, 100, 321}, K in {0, 1, 6, N-1} (stbmv/correctness.cpp:37-64);
tbsv K in {0, 1, 2, 10, 60, N-1} (stbsv/correctness.cpp:37-65).  tbsv uses
diagonally dominant bands (diagonal k + 1 + U[0,1], off-diagonal U[-1,1]) so
that the solve is well conditioned and the forward-error bound is meaningful.
Bounds: the K*eps bound of a length-K inner product (K = the band width).
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, NON_UNIT, TRANS, UNIT, UPPER, Oracle
from util import assert_bound, bits_equal, elems, eps, uniform, vec

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


def _band_dense(a, m, n, lo, hi, c0):
    D = np.zeros((m, n))
    for i in range(m):
        j0, j1 = max(0, i - lo), min(n, i + hi + 1)
        if j1 > j0:
            D[i, j0:j1] = a[i, c0 + j0 - i:c0 + j1 - i]
    return D


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("m,n,kl,ku", [(2, 2, 1, 1), (15, 20, 0, 5), (256, 256, 255, 255), (300, 120, 9, 40),
                                       (70, 90, 3, 0)])
@pytest.mark.parametrize("incx,incy,ld_diff", [(1, 1, 0), (3, -2, 3), (-1, 1, 0)])
def test_gbmv(C, dt, m, n, kl, ku, incx, incy, ld_diff, rng):
    lda = kl + ku + 1 + ld_diff
    a = uniform(rng, (m, lda), dt)
    # the gbmv wrapper does not re-base negative increments: it is given the element-0 address
    x, y = vec(rng, n, incx, dt), vec(rng, m, incy, dt)
    y0 = y.copy()
    yo = y.copy()
    Oracle(dt).gbmv(m, n, kl, ku, 1.2, 1.2, a, lda, x, incx, yo, incy)
    # element 0 of a negative-stride vector is its highest address
    base_x = x.ctypes.data + ((n - 1) * -incx * x.itemsize if incx < 0 else 0)
    base_y = y.ctypes.data + ((m - 1) * -incy * y.itemsize if incy < 0 else 0)
    fn(C, dt, "gbmv")(m, n, kl, ku, 1.2, a, lda, base_x, incx, 1.2, base_y, incy)
    D = np.abs(_band_dense(a.astype(np.float64), m, n, kl, ku, kl))
    mag = 1.2 * np.abs(elems(y0, m, incy)) + 1.2 * (D @ np.abs(elems(x, n, incx)))
    assert_bound(elems(y, m, incy), elems(yo, m, incy), mag, kl + ku + 1, dt, what="gbmv")


def _sbmv_cases():
    for n in (2, 100, 321):
        for k in (0, 1, 50, n - 1):
            if k <= n - 1:
                yield n, k
    return


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("n,k", sorted(set(_sbmv_cases())))
@pytest.mark.parametrize("ld_diff", [0, 3])
@pytest.mark.parametrize("incx,incy", [(1, 1), (3, -2), (1, -2), (3, 1)])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_sbmv(C, dt, uplo, n, k, ld_diff, incx, incy, beta, rng):
    A = uniform(rng, (n, k + 1 + ld_diff), dt)
    x, y = vec(rng, n, incx, dt), vec(rng, n, incy, dt)
    y0 = y.copy()
    yo = y.copy()
    o = Oracle(dt)
    if beta != 1.0:
        o.sbmv_scal_y(n, beta, yo, incy)
    o.sbmv(uplo, n, k, 2.4, A, A.shape[1], x, incx, yo, incy)
    fn(C, dt, "sbmv")(RM, uplo, n, k, 2.4, A, A.shape[1], x, incx, beta, y, incy)
    A64 = np.abs(A.astype(np.float64))
    T = _band_dense(A64, n, n, 0, k, 0) if uplo == UPPER else _band_dense(A64, n, n, k, 0, k)
    S = T + T.T
    mag = abs(beta) * np.abs(elems(y0, n, incy)) + 2.4 * (S @ np.abs(elems(x, n, incx)))
    assert_bound(elems(y, n, incy), elems(yo, n, incy), mag, 2 * k + 2, dt, what="sbmv")


def _tb_cases(ks):
    for n in (2, 100, 321):
        for k in ks + (n - 1,):
            if k <= n - 1:
                yield n, k


def _tri_band(rng, n, k, lda, uplo, dt):
    A = uniform(rng, (n, lda), dt)  # band columns 0..k, the rest padding
    dcol = 0 if uplo == UPPER else k
    A[:, dcol] = (k + 1 + rng.uniform(0, 1, n)).astype(dt)
    return A


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS],
                                                                      [NON_UNIT, UNIT])))
@pytest.mark.parametrize("n,k", sorted(set(_tb_cases((0, 1, 6)))))
@pytest.mark.parametrize("ld_diff", [0, 3])
@pytest.mark.parametrize("incx", [1, 3, -2])
def test_tbmv(C, dt, uplo, trans, diag, n, k, ld_diff, incx, rng):
    A = uniform(rng, (n, k + 1 + ld_diff), dt)
    x = vec(rng, n, incx, dt)
    x0 = x.copy()
    xo = x.copy()
    Oracle(dt).tbmv(uplo, trans, diag, n, k, xo, incx, A, A.shape[1])
    fn(C, dt, "tbmv")(RM, uplo, trans, diag, n, k, A, A.shape[1], x, incx)
    A64 = np.abs(A.astype(np.float64))
    T = _band_dense(A64, n, n, 0, k, 0) if uplo == UPPER else _band_dense(A64, n, n, k, 0, k)
    if diag == UNIT:
        np.fill_diagonal(T, 1.0)
    opT = T if trans == NO_TRANS else T.T
    mag = opT @ np.abs(elems(x0, n, incx))
    assert_bound(elems(x, n, incx), elems(xo, n, incx), mag, k + 1, dt, what="tbmv")
    if abs(incx) > 1:  # the gaps between strided elements are untouched
        mask = np.ones(x.size, bool)
        mask[np.arange(n) * abs(incx)] = False
        assert bits_equal(x[mask], x0[mask])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS],
                                                                      [NON_UNIT, UNIT])))
@pytest.mark.parametrize("n,k", sorted(set(_tb_cases((0, 1, 2, 10, 60)))))
@pytest.mark.parametrize("ld_diff", [0, 3])
@pytest.mark.parametrize("incx", [1, 3, -1])
def test_tbsv(C, dt, uplo, trans, diag, n, k, ld_diff, incx, rng):
    A = _tri_band(rng, n, k, k + 1 + ld_diff, uplo, dt)
    if diag == UNIT:  # unit diagonal: keep the off-diagonal sum below 1 per row
        A[:, :k + 1] /= (k + 1)
    x = vec(rng, n, incx, dt)
    xo = x.copy()
    Oracle(dt).tbsv(uplo, trans, diag, n, k, xo, incx, A, A.shape[1])
    fn(C, dt, "tbsv")(RM, uplo, trans, diag, n, k, A, A.shape[1], x, incx)
    want = elems(xo, n, incx).astype(np.float64)
    got = elems(x, n, incx).astype(np.float64)
    assert np.max(np.abs(got - want)) <= 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1), "tbsv"


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS],
                                                                      [NON_UNIT, UNIT])))
@pytest.mark.parametrize("n", [40, 512, 4097, 65539])
@pytest.mark.parametrize("k", [0, 3, 8, 16, 31, 32])
@pytest.mark.parametrize("incx", [1, -2])
def test_tbsv_blocked(C, dt, uplo, trans, diag, n, k, incx, rng):
    """Narrow bands (k <= 32): n = 40 takes the one-wave solve, 4097 and 65539
    the blocked solve (blocks of 64 / 256 rows, the last one partial), 512 the
    blocked solve for k <= 16 and trsv's block solver for k > 16 (faster
    there): every Uplo / Trans / Diag form, band widths at and between the
    window sizes (4, 8, 16, 32), a negative increment."""
    A = _tri_band(rng, n, k, k + 1, uplo, dt)
    if diag == UNIT:
        A[:, :k + 1] /= (k + 1)
    x = vec(rng, n, incx, dt)
    xo = x.copy()
    Oracle(dt).tbsv(uplo, trans, diag, n, k, xo, incx, A, A.shape[1])
    fn(C, dt, "tbsv")(RM, uplo, trans, diag, n, k, A, A.shape[1], x, incx)
    want = elems(xo, n, incx).astype(np.float64)
    got = elems(x, n, incx).astype(np.float64)
    assert np.max(np.abs(got - want)) <= 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1), "tbsv blocked"


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS],
                                                                      [NON_UNIT, UNIT])))
@pytest.mark.parametrize("n,k", [(52, 32), (53, 32), (108, 16), (109, 16), (231, 8), (232, 8), (511, 8),
                                 (512, 8), (1536, 32), (1537, 32)])
def test_tbsv_dispatch_edges(C, dt, uplo, trans, diag, n, k, rng):
    """Both sides of every point where tbsv changes solver: the one-wave solve
    / trsv's block solver at n = 53 (window 32), 109 (16) and 232 (8), the
    one-wave / blocked solve at n = 512, the block solver / blocked solve at
    n = 1537 (window 32)."""
    A = _tri_band(rng, n, k, k + 1, uplo, dt)
    if diag == UNIT:
        A[:, :k + 1] /= (k + 1)
    x = vec(rng, n, 1, dt)
    xo = x.copy()
    Oracle(dt).tbsv(uplo, trans, diag, n, k, xo, 1, A, A.shape[1])
    fn(C, dt, "tbsv")(RM, uplo, trans, diag, n, k, A, A.shape[1], x, 1)
    want = xo.astype(np.float64)
    assert np.max(np.abs(x.astype(np.float64) - want)) <= 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS],
                                                                      [NON_UNIT, UNIT])))
@pytest.mark.parametrize("n,k", [(700, 33), (700, 64), (700, 200), (4096, 64), (4096, 1024), (4096, 4095)])
@pytest.mark.parametrize("incx", [1, -2])
def test_tbsv_wide(C, dt, uplo, trans, diag, n, k, incx, rng):
    """Bands of more than 32 off-diagonals take trsv's block solver over the
    band; the shapes include the reference benchmark's N = 4096, K = N / 4
    and N / 64 (stbsv/bench.cpp:68-71) and a full triangle (K = N - 1)."""
    A = _tri_band(rng, n, k, k + 1, uplo, dt)
    if diag == UNIT:
        A[:, :k + 1] /= (k + 1)
    x = vec(rng, n, incx, dt)
    xo = x.copy()
    Oracle(dt).tbsv(uplo, trans, diag, n, k, xo, incx, A, A.shape[1])
    fn(C, dt, "tbsv")(RM, uplo, trans, diag, n, k, A, A.shape[1], x, incx)
    want = elems(xo, n, incx).astype(np.float64)
    got = elems(x, n, incx).astype(np.float64)
    assert np.max(np.abs(got - want)) <= 8 * (k + 2) * eps(dt) * (np.max(np.abs(want)) + 1), "tbsv wide"


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS])))
def test_banded_large_device(C, dt, uplo, trans):
    """n = 20000 bands of 64 (tbsv) and 200 (the products) on device buffers."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    tdt = torch.float32 if dt == np.float32 else torch.float64
    n, k = 20000, 200
    A = torch.empty((n, k + 1), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    x = torch.empty(n, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    y = torch.empty(n, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    An, xn, yn = A.cpu().numpy(), x.cpu().numpy(), y.cpu().numpy()
    o = Oracle(dt)
    yo = yn.copy()
    o.sbmv_scal_y(n, 0.5, yo, 1)
    o.sbmv(uplo, n, k, 1.5, An, k + 1, xn, 1, yo, 1)
    fn(C, dt, "sbmv")(RM, uplo, n, k, 1.5, A, k + 1, x, 1, 0.5, y, 1)
    torch.cuda.synchronize()
    assert_bound(y.cpu().numpy(), yo, 0.5 * np.abs(yn) + 1.5 * (2 * k + 1) * np.max(np.abs(xn)), 2 * k + 2, dt,
                 what="sbmv large")
    xo = xn.copy()
    o.tbmv(uplo, trans, NON_UNIT, n, k, xo, 1, An, k + 1)
    fn(C, dt, "tbmv")(RM, uplo, trans, NON_UNIT, n, k, A, k + 1, x, 1)
    torch.cuda.synchronize()
    assert_bound(x.cpu().numpy(), xo, (k + 1) * np.max(np.abs(xn)) * np.ones(n), k + 1, dt, what="tbmv large")
    ks = 64
    B = torch.empty((n, ks + 1), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    B[:, 0 if uplo == UPPER else ks] = ks + 1 + torch.empty(n, dtype=tdt, device="cuda").uniform_(0, 1, generator=g)
    b = torch.empty(n, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    Bn, bo = B.cpu().numpy(), b.cpu().numpy().copy()
    o.tbsv(uplo, trans, NON_UNIT, n, ks, bo, 1, Bn, ks + 1)
    fn(C, dt, "tbsv")(RM, uplo, trans, NON_UNIT, n, ks, B, ks + 1, b, 1)
    torch.cuda.synchronize()
    got = b.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - bo)) <= 8 * (ks + 2) * eps(dt) * (np.max(np.abs(bo)) + 1), "tbsv large"
