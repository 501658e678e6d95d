"""Level-3 parity: MFMA gemm / symm / syrk and the mscal / trmscal beta passes
behind the exo_* C-ABI, against the oracle.

Sweep of the reference's correctness tests: dims in {1, 7, 32, 64, 257}^3,
every trans / side / uplo combination, ld_diff in {0, 5}, alpha = 13,
beta = 18 (test/level3/gemm/correctness.cpp:45-64, symm/correctness.cpp:43-62,
syrk/correctness.cpp:39-55), through exoblas_amd.cblas, which mirrors the
wrappers (mscal / trmscal for beta, then the product).  The full-size cases
(config 1: sgemm NN 512^3 seed 0 alpha 1.5 beta 0.5; 4096 / 8192 products)
compare against a float64 GEMM within the K*eps bound.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import LEFT, LOWER, NO_TRANS, RIGHT, TRANS, UPPER, Oracle
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


def padded(rng, rows, cols, ld_diff, dt):
    return uniform(rng, (rows, cols + ld_diff), dt)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("ta,tb", list(itertools.product([NO_TRANS, TRANS], [NO_TRANS, TRANS])))
@pytest.mark.parametrize("M,N,K", list(itertools.product(DIMS, DIMS, DIMS)))
def test_gemm(C, dt, ta, tb, M, N, K, rng):
    ld_diff = 5 if (M + N + K) % 2 else 0
    A = padded(rng, M if ta == NO_TRANS else K, K if ta == NO_TRANS else M, ld_diff, dt)
    B = padded(rng, K if tb == NO_TRANS else N, N if tb == NO_TRANS else K, ld_diff, dt)
    Cm = padded(rng, M, N, ld_diff, dt)
    lda, ldb, ldc = A.shape[1], B.shape[1], Cm.shape[1]
    Co = Cm.copy()
    o = Oracle(dt)
    o.mscal(M, N, 18.0, Co, ldc)
    o.gemm(ta, tb, M, N, K, 13.0, A, lda, B, ldb, Co, ldc)
    C0 = Cm.copy()
    fn(C, dt, "gemm")(RM, ta, tb, M, N, K, 13.0, A, lda, B, ldb, 18.0, Cm, ldc)
    opA = np.abs(A[:, :K] if ta == NO_TRANS else A[:, :M].T).astype(np.float64)
    opB = np.abs(B[:, :N] if tb == NO_TRANS else B[:, :K].T).astype(np.float64)
    mag = np.zeros(Cm.shape)
    mag[:, :N] = 18.0 * np.abs(C0[:, :N]) + 13.0 * (opA @ opB)
    assert_bound(Cm, Co, mag, K, dt, what="gemm")
    assert bits_equal(Cm[:, N:], C0[:, N:])  # padding untouched


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("M,N", list(itertools.product(DIMS, DIMS)))
@pytest.mark.parametrize("ld_diff", [0, 5])
def test_symm(C, dt, side, uplo, M, N, ld_diff, rng):
    na = M if side == LEFT else N
    A = padded(rng, na, na, ld_diff, dt)
    B = padded(rng, M, N, ld_diff, dt)
    Cm = padded(rng, M, N, ld_diff, dt)
    Co = Cm.copy()
    o = Oracle(dt)
    o.mscal(M, N, 18.0, Co, Cm.shape[1])
    o.symm(side, uplo, M, N, 13.0, A, A.shape[1], B, B.shape[1], Co, Cm.shape[1])
    C0 = Cm.copy()
    fn(C, dt, "symm")(RM, side, uplo, M, N, 13.0, A, A.shape[1], B, B.shape[1], 18.0, Cm, Cm.shape[1])
    T = np.tril(A[:, :na]) if uplo == LOWER else np.triu(A[:, :na])
    S = np.abs((T + T.T - np.diag(np.diag(T))).astype(np.float64))
    Bb = np.abs(B[:, :N].astype(np.float64))
    mag = np.zeros(Cm.shape)
    mag[:, :N] = 18.0 * np.abs(C0[:, :N]) + 13.0 * (S @ Bb if side == LEFT else Bb @ S)
    assert_bound(Cm, Co, mag, na, dt, what="symm")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("N,K", list(itertools.product(DIMS, DIMS)))
@pytest.mark.parametrize("ld_diff", [0, 5])
def test_syrk(C, dt, uplo, trans, N, K, ld_diff, rng):
    A = padded(rng, N if trans == NO_TRANS else K, K if trans == NO_TRANS else N, ld_diff, dt)
    Cm = padded(rng, N, N, ld_diff, dt)
    Co = Cm.copy()
    o = Oracle(dt)
    o.trmscal(uplo, N, 18.0, Co, Cm.shape[1])
    o.syrk(uplo, trans, N, K, 13.0, A, A.shape[1], Co, Cm.shape[1])
    C0 = Cm.copy()
    fn(C, dt, "syrk")(RM, uplo, trans, N, K, 13.0, A, A.shape[1], 18.0, Cm, Cm.shape[1])
    Aa = np.abs(A[:, :K] if trans == NO_TRANS else A[:, :N].T).astype(np.float64)  # N x K
    mag = np.zeros(Cm.shape)
    mag[:, :N] = 18.0 * np.abs(C0[:, :N]) + 13.0 * (Aa @ Aa.T)
    assert_bound(Cm, Co, mag, K, dt, what="syrk")
    mask = np.tril(np.ones((N, N), bool)) if uplo == LOWER else np.triu(np.ones((N, N), bool))
    assert bits_equal(Cm[:, :N][~mask], C0[:, :N][~mask])  # other triangle untouched


def _f64_gemm(torch, a, b):
    return (a.double() @ b.double())


def test_sgemm_config1(C):
    """sgemm NN 512^3, A, B, C ~ U[-1,1] from seed 0, alpha 1.5, beta 0.5,
    against the oracle and against float64."""
    rng = np.random.default_rng(0)
    n = 512
    A = uniform(rng, (n, n), np.float32)
    B = uniform(rng, (n, n), np.float32)
    Cm = uniform(rng, (n, n), np.float32)
    C0 = Cm.copy()
    Co = Cm.copy()
    o = Oracle(np.float32)
    o.mscal(n, n, 0.5, Co, n)
    o.gemm(NO_TRANS, NO_TRANS, n, n, n, 1.5, A, n, B, n, Co, n)
    C.exo_sgemm(RM, NO_TRANS, NO_TRANS, n, n, n, 1.5, A, n, B, n, 0.5, Cm, n)
    mag = 0.5 * np.abs(C0) + 1.5 * (np.abs(A).astype(np.float64) @ np.abs(B))
    assert_bound(Cm, Co, mag, n, np.float32, what="sgemm 512 vs oracle")
    truth = 0.5 * C0.astype(np.float64) + 1.5 * (A.astype(np.float64) @ B.astype(np.float64))
    assert_bound(Cm, truth, mag, n, np.float32, c=1.0, what="sgemm 512 vs fp64")


@pytest.mark.parametrize("dt,n", [(np.float32, 8192), (np.float64, 4096)])
def test_gemm_full_size_device(C, dt, n):
    """Device-resident NN product at full size against a float64 GEMM."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    tdt = torch.float32 if dt == np.float32 else torch.float64
    A = torch.empty((n, n), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    B = torch.empty((n, n), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    Cm = torch.empty((n, n), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    C0 = Cm.clone()
    fn(C, dt, "gemm")(RM, NO_TRANS, NO_TRANS, n, n, n, 1.5, A, n, B, n, 0.5, Cm, n)
    torch.cuda.synchronize()
    truth = 0.5 * C0.double() + 1.5 * (A.double() @ B.double())
    mag = 0.5 * C0.double().abs() + 1.5 * (A.double().abs() @ B.double().abs())
    eps = float(np.finfo(dt).eps)
    err = (Cm.double() - truth).abs()
    ratio = (err / (mag * (n + 2) * eps)).max().item()
    assert ratio <= 1.0, ratio
    del A, B, Cm, C0, truth, mag, err
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo", [UPPER, LOWER])
def test_mscal_trmscal(C, dt, uplo, rng):
    from exoblas_amd import lib as L
    import ctypes
    so = L.load()
    p = "s" if dt == np.float32 else "d"
    T = ctypes.c_float if dt == np.float32 else ctypes.c_double
    for (M, N, ld) in [(7, 5, 5), (33, 257, 262), (300, 300, 300), (2, 1, 3)]:
        A = uniform(rng, (M, ld), dt)
        Ao = A.copy()
        Oracle(dt).mscal(M, N, 0.3, Ao, ld)
        getattr(so, f"exo_{p}mscal_rm_stride_1")(None, M, N, ctypes.byref(T(0.3)), L.win2(A.ctypes.data, ld))
        L.check()
        assert bits_equal(A, Ao)
        if M == N:
            Ao = A.copy()
            Oracle(dt).trmscal(uplo, N, -1.7, Ao, ld)
            getattr(so, f"exo_{p}trmscal_rm_stride_1")(None, uplo, N, ctypes.byref(T(-1.7)),
                                                       L.win2(A.ctypes.data, ld))
            L.check()
            assert bits_equal(A, Ao)


# Shapes that reach the largest tile configuration (at least two workgroup
# tiles per CU) for each form: FP32 256 x 256 tiles, FP64 128 x 256.
HUGE = {np.float32: dict(mn=(5800, 5784), k=700, syrk=(8192, 260)),
        np.float64: dict(mn=(4100, 4104), k=300, syrk=(6000, 200))}


def _dev(torch, rows, cols, dt, g):
    tdt = torch.float32 if dt == np.float32 else torch.float64
    return torch.empty((rows, cols), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)


def _check_dev(torch, got, truth, mag, K, dt):
    eps = float(np.finfo(dt).eps)
    # mag is 0 only where C0 is exactly 0 and nothing is added (err is 0 there)
    ratio = ((got.double() - truth).abs() / (mag * (K + 2) * eps).clamp_min(1e-300)).max().item()
    assert ratio <= 1.0, ratio


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("ta,tb,ld_diff", [(NO_TRANS, NO_TRANS, 0), (NO_TRANS, TRANS, 0), (TRANS, NO_TRANS, 0),
                                           (TRANS, TRANS, 0), (NO_TRANS, NO_TRANS, 3)])
def test_gemm_huge_tiles(C, dt, ta, tb, ld_diff):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    (M, N), K = HUGE[dt]["mn"], HUGE[dt]["k"]
    A = _dev(torch, M if ta == NO_TRANS else K, (K if ta == NO_TRANS else M) + ld_diff, dt, g)
    B = _dev(torch, K if tb == NO_TRANS else N, (N if tb == NO_TRANS else K) + ld_diff, dt, g)
    Cm = _dev(torch, M, N + ld_diff, dt, g)
    C0 = Cm.clone()
    fn(C, dt, "gemm")(RM, ta, tb, M, N, K, 1.5, A, A.shape[1], B, B.shape[1], 0.5, Cm, Cm.shape[1])
    torch.cuda.synchronize()
    opA = A[:, :K] if ta == NO_TRANS else A[:, :M].T
    opB = B[:, :N] if tb == NO_TRANS else B[:, :K].T
    truth = 0.5 * C0[:, :N].double() + 1.5 * (opA.double() @ opB.double())
    mag = 0.5 * C0[:, :N].double().abs() + 1.5 * (opA.double().abs() @ opB.double().abs())
    _check_dev(torch, Cm[:, :N], truth, mag, K, dt)
    assert torch.equal(Cm[:, N:], C0[:, N:])  # padding columns untouched


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo", list(itertools.product([LEFT, RIGHT], [UPPER, LOWER])))
def test_symm_huge_tiles(C, dt, side, uplo):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    M, N = HUGE[dt]["mn"]
    na = M if side == LEFT else N
    A = _dev(torch, na, na, dt, g)
    B = _dev(torch, M, N, dt, g)
    Cm = _dev(torch, M, N, dt, g)
    C0 = Cm.clone()
    fn(C, dt, "symm")(RM, side, uplo, M, N, 1.5, A, na, B, N, 0.5, Cm, N)
    torch.cuda.synchronize()
    T = A.double().tril() if uplo == LOWER else A.double().triu()
    S = T + T.T - torch.diag(torch.diag(T))
    prod = S @ B.double() if side == LEFT else B.double() @ S
    mag_p = S.abs() @ B.double().abs() if side == LEFT else B.double().abs() @ S.abs()
    _check_dev(torch, Cm, 0.5 * C0.double() + 1.5 * prod, 0.5 * C0.double().abs() + 1.5 * mag_p, na, dt)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS])))
def test_syrk_huge_tiles(C, dt, uplo, trans):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    N, K = HUGE[dt]["syrk"]
    A = _dev(torch, N if trans == NO_TRANS else K, K if trans == NO_TRANS else N, dt, g)
    Cm = _dev(torch, N, N, dt, g)
    C0 = Cm.clone()
    fn(C, dt, "syrk")(RM, uplo, trans, N, K, 1.5, A, A.shape[1], 0.5, Cm, N)
    torch.cuda.synchronize()
    X = (A if trans == NO_TRANS else A.T).double()
    mask = torch.ones((N, N), dtype=torch.bool, device="cuda")
    mask = mask.tril() if uplo == LOWER else mask.triu()
    truth = torch.where(mask, 0.5 * C0.double() + 1.5 * (X @ X.T), C0.double())
    mag = torch.where(mask, 0.5 * C0.double().abs() + 1.5 * (X.abs() @ X.abs().T), C0.double().abs())
    _check_dev(torch, Cm, truth, mag, K, dt)
    assert torch.equal(Cm[~mask], C0[~mask])  # other triangle untouched
