"""Split-precision sgemm / ssyrk (FP32 operands as BF16 pieces on the BF16
MFMA) against the oracle and against float64.

Tolerances are written per form, from the representation error of the split:
  terms = 3 (two pieces):   |x - h - m| <= 2^-17 |x|, the dropped m*m term
                            <= 2^-18 |ab|: per product ~ 2^-16 |a b|
  terms = 4 (two pieces + m*m): the same bound (h*r and r*h remain)
  terms = 6 (three pieces): per product ~ 2^-24 |a b|, the FP32 level
on top of the FP32 accumulation bound (K+2) eps_32 sum|a b|.
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, TRANS, UPPER, Oracle
from util import uniform

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
PER_PRODUCT = {3: 2.0 ** -16, 4: 2.0 ** -16, 6: 2.0 ** -23}


@pytest.fixture(scope="module")
def so():
    from exoblas_amd import lib
    return lib.load()


def check(got, want, mag, K, terms, what):
    tol = 4 * (K + 2) * (EPS32 + PER_PRODUCT[terms]) * mag + 1e-30
    err = np.abs(np.asarray(got, np.float64) - want)
    bad = err > tol
    assert not bad.any(), f"{what}: {int(bad.sum())} out of bound, worst ratio {float((err / tol).max()):.3g}"


@pytest.mark.parametrize("terms", [3, 4, 6])
@pytest.mark.parametrize("ta,tb", list(itertools.product([NO_TRANS, TRANS], [NO_TRANS, TRANS])))
@pytest.mark.parametrize("M,N,K", [(1, 7, 32), (7, 257, 64), (257, 64, 257), (130, 129, 33), (256, 256, 256)])
def test_gemm_split_vs_oracle(so, terms, ta, tb, M, N, K, rng):
    from exoblas_amd import lib as L
    A = uniform(rng, (M, K) if ta == NO_TRANS else (K, M), np.float32)
    B = uniform(rng, (K, N) if tb == NO_TRANS else (N, K), np.float32)
    Cm = uniform(rng, (M, N), np.float32)
    Co = Cm.copy()
    Oracle(np.float32).gemm(ta, tb, M, N, K, 1.5, A, A.shape[1], B, B.shape[1], Co, N)
    C0 = Cm.copy()
    a = L.win2(A.ctypes.data, A.shape[1])
    b = L.win2(B.ctypes.data, B.shape[1])
    st = so.exoblas_sgemm_split(None, ta, tb, M, N, K, ctypes.byref(ctypes.c_float(1.5)), a, b, a, b,
                                L.win2(Cm.ctypes.data, N), terms)
    assert st == 0, L.last_error()
    opA = np.abs(A if ta == NO_TRANS else A.T).astype(np.float64)
    opB = np.abs(B if tb == NO_TRANS else B.T).astype(np.float64)
    check(Cm, Co, np.abs(C0) + 1.5 * (opA @ opB), K, terms, "gemm_split")


@pytest.mark.parametrize("terms", [3, 4, 6])
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("N,K", [(7, 33), (257, 64), (300, 129)])
def test_syrk_split_vs_oracle(so, terms, uplo, trans, N, K, rng):
    from exoblas_amd import lib as L
    A = uniform(rng, (N, K) if trans == NO_TRANS else (K, N), np.float32)
    Cm = uniform(rng, (N, N), np.float32)
    Co = Cm.copy()
    Oracle(np.float32).syrk(uplo, trans, N, K, 2.0, A, A.shape[1], Co, N)
    C0 = Cm.copy()
    a = L.win2(A.ctypes.data, A.shape[1])
    st = so.exoblas_ssyrk_split(None, uplo, trans, N, K, ctypes.byref(ctypes.c_float(2.0)), a, a,
                                L.win2(Cm.ctypes.data, N), terms)
    assert st == 0, L.last_error()
    Aa = np.abs(A if trans == NO_TRANS else A.T).astype(np.float64)
    check(Cm, Co, np.abs(C0) + 2.0 * (Aa @ Aa.T), K, terms, "syrk_split")
    mask = np.tril(np.ones((N, N), bool)) if uplo == LOWER else np.triu(np.ones((N, N), bool))
    assert np.array_equal(Cm[~mask], C0[~mask])


@pytest.mark.parametrize("n", [2048, 8192])
def test_split_error_vs_plain_fp32(so, n):
    """n^3 NN on the uniform and the row-scaled (2^U[-8,8]) sets: error of
    plain FP32 MFMA and the 3-, 4- and 6-product splits against float64,
    normalised by sum|a b|.  The truncation error of the 2-piece forms falls
    relative to sum|a b| as 1/sqrt(K) while plain FP32's accumulation error
    grows with K: at K = 8192 the 4- and 6-product forms are at most plain
    FP32's error, at K = 2048 both 2-piece forms stay within their
    per-product bound.  At n = 2048 plain sgemm splits K four ways (64 large
    tiles for 256 CUs), which shortens its accumulation chains, and the
    6-product form stays within twice its error."""
    import torch
    from exoblas_amd import lib as L
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    stream = torch.cuda.current_stream().cuda_stream
    for scaled in (False, True):
        A = torch.empty((n, n), device="cuda").uniform_(-1, 1, generator=g)
        B = torch.empty((n, n), device="cuda").uniform_(-1, 1, generator=g)
        if scaled:
            A *= torch.exp2(torch.empty((n, 1), device="cuda").uniform_(-8, 8, generator=g))
            B *= torch.exp2(torch.empty((n, 1), device="cuda").uniform_(-8, 8, generator=g))
        truth = A.double() @ B.double()
        mag = A.double().abs() @ B.double().abs()
        errs = {}
        Cm = torch.zeros((n, n), device="cuda")
        for form in ("plain", 3, 4, 6):
            Cm.zero_()
            a, b, c = L.win2(A.data_ptr(), n), L.win2(B.data_ptr(), n), L.win2(Cm.data_ptr(), n)
            one = ctypes.byref(ctypes.c_float(1.0))
            if form == "plain":
                st = so.exoblas_sgemm_stride_1_async(stream, NO_TRANS, NO_TRANS, n, n, n, one, a, b, a, b, c)
            else:
                st = so.exoblas_sgemm_split_async(stream, NO_TRANS, NO_TRANS, n, n, n, one, a, b, a, b, c, form)
            assert st == 0
            torch.cuda.synchronize()
            errs[form] = ((Cm.double() - truth).abs() / mag).max().item()
        if n >= 8192:
            assert errs[6] <= errs["plain"], errs
            assert errs[4] <= errs["plain"], errs
        else:
            assert errs[6] <= 2 * errs["plain"], errs
        assert max(errs[3], errs[4]) <= 4 * n * (EPS32 + PER_PRODUCT[3]), errs
        del A, B, Cm, truth, mag
        torch.cuda.empty_cache()


@pytest.mark.parametrize("update", [16, 32])
@pytest.mark.parametrize("refine", [0, 1])
@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS], [131, 132])))
@pytest.mark.parametrize("n", [65, 300, 1100])
def test_trsv_mixed(so, update, refine, uplo, trans, diag, n, rng):
    """Against a float64 solve of the same FP32 matrix: one refinement step
    brings the solution to FP32 rounding of the exact one (|err| <= 4 eps_32
    |x|); without refinement the BF16 off-diagonal updates leave ~2^-8 of the
    off-diagonal contribution."""
    from exoblas_amd import lib as L
    A = (rng.uniform(-1, 1, (n, n)) / n).astype(np.float32)
    A[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    T = np.tril(A) if uplo == LOWER else np.triu(A)
    T = T.astype(np.float64)
    if diag == 132:
        np.fill_diagonal(T, 1.0)
    opT = T if trans == NO_TRANS else T.T
    exact = np.linalg.solve(opT, b.astype(np.float64))
    x = b.copy()
    st = so.exoblas_strsv_mixed(None, uplo, trans, diag, n, L.win1(x.ctypes.data, 1), L.win2(A.ctypes.data, n),
                                update, refine)
    assert st == 0, L.last_error()
    err = np.abs(x.astype(np.float64) - exact)
    if refine:
        # FP32 rounding of the refined solution, plus the correction's own
        # error (2^-8 of |d| ~ eps |x| for BF16-stored updates, n eps of it
        # for FP32 ones), which is normwise: it only shows on entries far
        # below max |x|
        corr = (2.0 ** -8 if update == 16 else n * EPS32) * 4 * EPS32 * np.abs(exact).max()
        tol = 4 * EPS32 * np.abs(exact) + corr + 1e-30
    else:
        off = np.abs(opT - np.diag(np.diag(opT))) @ np.abs(exact)
        per = 2.0 ** -8 if update == 16 else 4 * n * EPS32
        tol = 4 * EPS32 * np.abs(exact) + per * off * (3.0 if diag == 132 else 3.0 / n) + 1e-30
    assert np.all(err <= tol), float((err / tol).max())


@pytest.mark.parametrize("uplo,trans", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS])))
@pytest.mark.parametrize("incx", [2, -3])
def test_trsv_mixed_strided(so, uplo, trans, incx, rng):
    """The fused refined solve (FP32 updates, one refinement) reads b and
    writes x through the caller's stride; entries between the strided
    elements stay untouched."""
    from exoblas_amd import lib as L
    n = 1100
    A = (rng.uniform(-1, 1, (n, n)) / n).astype(np.float32)
    A[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    T = (np.tril(A) if uplo == LOWER else np.triu(A)).astype(np.float64)
    opT = T if trans == NO_TRANS else T.T
    exact = np.linalg.solve(opT, b.astype(np.float64))
    step = abs(incx)
    buf = rng.uniform(-1, 1, (n - 1) * step + 1).astype(np.float32)
    view = buf[::step] if incx > 0 else buf[::step][::-1]
    view[:] = b
    before = buf.copy()
    base = buf.ctypes.data if incx > 0 else buf.ctypes.data + (n - 1) * step * 4
    st = so.exoblas_strsv_mixed(None, uplo, trans, 131, n, L.win1(base, incx), L.win2(A.ctypes.data, n), 32, 1)
    assert st == 0, L.last_error()
    x = (buf[::step] if incx > 0 else buf[::step][::-1]).astype(np.float64)
    assert np.all(np.abs(x - exact) <= 4 * EPS32 * np.abs(exact) + 1e-30)
    mask = np.ones(buf.size, bool)
    mask[::step] = False
    assert np.array_equal(buf[mask], before[mask])


@pytest.mark.parametrize("uplo,trans", [(LOWER, NO_TRANS), (UPPER, TRANS), (UPPER, NO_TRANS)])
def test_trsv_mixed_packed(so, uplo, trans, rng):
    """The BF16 copy packed once and reused: two solves with different
    right-hand sides through exoblas_strsv_mixed_packed_async give the same
    bits as exoblas_strsv_mixed_async with update 16 (same kernels, the same
    copy), and a bad argument is rejected with its status."""
    import torch
    from exoblas_amd import lib as L
    n = 700
    A = (rng.uniform(-1, 1, (n, n)) / n).astype(np.float32)
    A[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(np.float32)
    dA = torch.from_numpy(A).cuda()
    packed = torch.empty(n * n, dtype=torch.bfloat16, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    wA = L.win2(dA.data_ptr(), n)
    assert so.exoblas_strsv_mixed_pack_async(sh, n, wA, packed.data_ptr()) == 0, L.last_error()
    for refine in (1, 0):
        b = torch.from_numpy(rng.uniform(-1, 1, n).astype(np.float32)).cuda()
        x1, x2 = b.clone(), b.clone()
        assert so.exoblas_strsv_mixed_packed_async(sh, uplo, trans, 131, n, L.win1(x1.data_ptr(), 1), wA,
                                                   packed.data_ptr(), refine) == 0, L.last_error()
        assert so.exoblas_strsv_mixed_async(sh, uplo, trans, 131, n, L.win1(x2.data_ptr(), 1), wA, 16, refine) == 0
        torch.cuda.synchronize()
        assert torch.equal(x1, x2), f"packed solve differs (refine {refine})"
    x = torch.zeros(n, device="cuda")
    assert so.exoblas_strsv_mixed_packed_async(sh, uplo, trans, 131, n, L.win1(x.data_ptr(), 1), wA,
                                               packed.data_ptr(), -1) != 0
    assert so.exoblas_strsv_mixed_packed_async(sh, 120, trans, 131, n, L.win1(x.data_ptr(), 1), wA,
                                               packed.data_ptr(), 1) != 0
