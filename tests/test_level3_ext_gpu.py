"""Parity of syr2k and strmm (the reference's exo_syr2k and exo_trmm library
targets) behind the exo_* C-ABI, against the oracle.

syr2k: the reference's sweep (test/level3/syr2k/correctness.cpp:43-64): N, K
in {1, 7, 32, 64, 257}, both Uplo and Trans, alpha = 13, beta = 18, ld
differences of 0 and 5, through exoblas_amd.cblas, which mirrors the wrapper
(trmscal for beta, then syr2k_rm_stride_1).  strmm: the reference's sweep
(test/level3/strmm/correctness.cpp:44-75): M in {4, 101, 321}, N in {8, 15,
564}, Left side, both Uplo / TransA / Diag, alpha in {0, 1, 1.4}, lda and ldb
differences of 0 and 5.  Bounds: the K*eps forward-error bound of a length-K
inner product (K = 2K for syr2k's two products, K = M for trmm).  Full-size
device cases compare against float64 products.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import LEFT, LOWER, NO_TRANS, NON_UNIT, RIGHT, TRANS, UNIT, UPPER, Oracle
from util import assert_bound, bits_equal, uniform

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
RM = 101
DIMS = [1, 7, 32, 64, 257]


@pytest.fixture(scope="module")
def C():
    from exoblas_amd import cblas, lib
    lib.load()
    return cblas


def fn(C, dt, name):
    return getattr(C, ("exo_s" if dt == np.float32 else "exo_d") + name)


def _tri_mask(N, uplo):
    return np.tril(np.ones((N, N), bool)) if uplo == LOWER else np.triu(np.ones((N, N), bool))


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("N,K", list(itertools.product(DIMS, DIMS)))
@pytest.mark.parametrize("lda_d,ldb_d,ldc_d", [(0, 0, 0), (5, 5, 5), (0, 5, 0), (5, 0, 5)])
def test_syr2k(C, dt, uplo, trans, N, K, lda_d, ldb_d, ldc_d, rng):
    rows, cols = (N, K) if trans == NO_TRANS else (K, N)
    A = uniform(rng, (rows, cols + lda_d), dt)
    B = uniform(rng, (rows, cols + ldb_d), dt)
    Cm = uniform(rng, (N, N + ldc_d), dt)
    C0 = Cm.copy()
    Co = Cm.copy()
    o = Oracle(dt)
    o.trmscal(uplo, N, 18.0, Co, Cm.shape[1])
    o.syr2k(uplo, trans, N, K, 13.0, A, A.shape[1], B, B.shape[1], Co, Cm.shape[1])
    fn(C, dt, "syr2k")(RM, uplo, trans, N, K, 13.0, A, A.shape[1], B, B.shape[1], 18.0, Cm, Cm.shape[1])
    Aa = np.abs(A[:, :cols] if trans == NO_TRANS else A[:, :cols].T).astype(np.float64)  # N x K
    Ba = np.abs(B[:, :cols] if trans == NO_TRANS else B[:, :cols].T).astype(np.float64)
    mag = np.abs(C0).astype(np.float64)
    mag[:, :N] = 18.0 * np.abs(C0[:, :N]) + 13.0 * (Aa @ Ba.T + Ba @ Aa.T)
    assert_bound(Cm, Co, mag, 2 * K, dt, what="syr2k")
    mask = _tri_mask(N, uplo)
    assert bits_equal(Cm[:, :N][~mask], C0[:, :N][~mask])  # other triangle untouched
    assert bits_equal(Cm[:, N:], C0[:, N:])  # padding untouched


@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("diag", [NON_UNIT, UNIT])
@pytest.mark.parametrize("M,N", list(itertools.product([4, 101, 321], [8, 15, 564])))
@pytest.mark.parametrize("alpha", [0.0, 1.0, 1.4])
@pytest.mark.parametrize("lda_d,ldb_d", [(0, 0), (5, 5), (0, 5), (5, 0)])
def test_strmm(C, uplo, trans, diag, M, N, alpha, lda_d, ldb_d, rng):
    dt = np.float32
    A = uniform(rng, (M, M + lda_d), dt)
    B = uniform(rng, (M, N + ldb_d), dt)
    B0 = B.copy()
    Bo = B.copy()
    Oracle(dt).trmm(uplo, trans, diag, M, N, alpha, A, A.shape[1], Bo, B.shape[1])
    C.exo_strmm(RM, LEFT, uplo, trans, diag, M, N, alpha, A, A.shape[1], B, B.shape[1])
    T = np.abs(np.tril(A[:, :M]) if uplo == LOWER else np.triu(A[:, :M])).astype(np.float64)
    if diag == UNIT:
        np.fill_diagonal(T, 1.0)
    opT = T if trans == NO_TRANS else T.T
    mag = np.abs(B0).astype(np.float64)
    mag[:, :N] = abs(alpha) * (opT @ np.abs(B0[:, :N]).astype(np.float64))
    assert_bound(B, Bo, mag, M, dt, what="strmm")
    assert bits_equal(B[:, N:], B0[:, N:])  # padding untouched


def test_strmm_right_side_makes_no_call(C, rng):
    # the reference wrapper's Right branch is empty (exo_strmm.h:44-55)
    A = uniform(rng, (15, 15), np.float32)
    B = uniform(rng, (7, 15), np.float32)
    B0 = B.copy()
    C.exo_strmm(RM, RIGHT, LOWER, NO_TRANS, NON_UNIT, 7, 15, 1.4, A, 15, B, 15)
    assert bits_equal(B, B0)


def _dev(torch, rows, cols, dt, g):
    tdt = torch.float32 if dt == np.float32 else torch.float64
    return torch.empty((rows, cols), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)


def _check_dev(torch, got, truth, mag, K, dt):
    eps = float(np.finfo(dt).eps)
    ratio = ((got.double() - truth).abs() / (mag * (K + 2) * eps).clamp_min(1e-300)).max().item()
    assert ratio <= 1.0, ratio


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product([LOWER, UPPER], [NO_TRANS, TRANS])))
def test_syr2k_full_size_device(C, dt, uplo, trans):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    N, K = 4096, 1024
    shape = (N, K) if trans == NO_TRANS else (K, N)
    A, B = _dev(torch, *shape, dt, g), _dev(torch, *shape, dt, g)
    Cm = _dev(torch, N, N, dt, g)
    C0 = Cm.clone()
    fn(C, dt, "syr2k")(RM, uplo, trans, N, K, 1.5, A, shape[1], B, shape[1], 0.5, Cm, N)
    torch.cuda.synchronize()
    X = (A if trans == NO_TRANS else A.T).double()
    Y = (B if trans == NO_TRANS else B.T).double()
    mask = torch.ones((N, N), dtype=torch.bool, device="cuda")
    mask = mask.tril() if uplo == LOWER else mask.triu()
    truth = torch.where(mask, 0.5 * C0.double() + 1.5 * (X @ Y.T + Y @ X.T), C0.double())
    mag = torch.where(mask, 0.5 * C0.double().abs() + 1.5 * (X.abs() @ Y.abs().T + Y.abs() @ X.abs().T),
                      C0.double().abs())
    _check_dev(torch, Cm, truth, mag, 2 * K, dt)
    assert torch.equal(Cm[~mask], C0[~mask])


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([LOWER, UPPER], [NO_TRANS, TRANS],
                                                                      [NON_UNIT, UNIT])))
@pytest.mark.parametrize("M,N", [(4096, 4096), (5000, 3001)])
def test_strmm_full_size_device(C, uplo, trans, diag, M, N):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    A = _dev(torch, M, M, np.float32, g)
    B = _dev(torch, M, N, np.float32, g)
    B0 = B.clone()
    C.exo_strmm(RM, LEFT, uplo, trans, diag, M, N, 0.75, A, M, B, N)
    torch.cuda.synchronize()
    T = (A.tril() if uplo == LOWER else A.triu()).double()
    if diag == UNIT:
        T.fill_diagonal_(1.0)
    opT = T if trans == NO_TRANS else T.T
    truth = 0.75 * (opT @ B0.double())
    mag = 0.75 * (opT.abs() @ B0.double().abs())
    _check_dev(torch, B, truth, mag, M, np.float32)
