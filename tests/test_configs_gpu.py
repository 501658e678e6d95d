"""Parity at the benchmark shapes: the exact operands the benchmark times.

* triangular solves: s/d trsv lower, no-trans, non-unit at n = 4096 and 8192
  on the diagonally dominant matrices (off-diagonal U[-1,1]/n, diagonal
  n + U[0,1]), b ~ U[-1,1]; against the oracle (strict loop nest of
  src/level2/trsv.py:10-52) and a float64 solve;
* s/d syrk lower-N at N = 4096, K = 1024 and s/d symm left-lower at
  M = N = 4096 (their own tile plans: 128 x 128 tiles in two rounds plus a
  K-split round for syrk, the large tiles for symm); against a float64
  product within the (K + 2) eps bound;
* split ssyrk lower-N at N = 8192, K = 2048 on the uniform set and on rows
  scaled by 2^U[-8,8]: both error metrics (max |C - C64| / (|A||A|^T) and
  max |C - C64| / |C64|) stated for plain FP32 and each split form;
* the mixed strsv at n = 8192 with one refinement step, against the float64
  solve of the same FP32 matrix.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, NON_UNIT, Oracle
from util import assert_bound

pytestmark = pytest.mark.gpu

RM = 101
EPS32 = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def so():
    from exoblas_amd import lib
    return lib.load()


def dd_matrix(rng, n, dt):
    A = (rng.uniform(-1, 1, (n, n)) / n).astype(dt)
    A[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(dt)
    return A


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [4096, 8192])
def test_trsv_config_shapes(so, dt, n):
    import scipy.linalg
    import torch
    from exoblas_amd import lib as L
    rng = np.random.default_rng(0)
    A = dd_matrix(rng, n, dt)
    b = rng.uniform(-1, 1, n).astype(dt)
    xo = b.copy()
    Oracle(dt).trsv(LOWER, NO_TRANS, NON_UNIT, n, xo, 1, A, n)
    dA, dx = torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda()
    p = "s" if dt == np.float32 else "d"
    stream = torch.cuda.current_stream().cuda_stream
    st = getattr(so, f"exoblas_{p}trsv_rm_stride_1_async")(stream, LOWER, NO_TRANS, NON_UNIT, n,
                                                             L.win1(dx.data_ptr(), 1), L.win2(dA.data_ptr(), n))
    assert st == 0, L.last_error()
    x = dx.cpu().numpy()
    assert_bound(x, xo, np.abs(xo) + 1, n, dt, what=f"{p}trsv {n} vs oracle")
    truth = scipy.linalg.solve_triangular(np.tril(A).astype(np.float64), b.astype(np.float64), lower=True)
    # forward error of a substitution on a diagonally dominant matrix: a few
    # ulps of |x| (the off-diagonal part is ~1/n of the diagonal)
    assert_bound(x, truth, np.abs(truth), 8, dt, what=f"{p}trsv {n} vs fp64")
    # and a second, repeated solve on the same stream gives the same bits
    dx.copy_(torch.from_numpy(b))
    st = getattr(so, f"exoblas_{p}trsv_rm_stride_1_async")(stream, LOWER, NO_TRANS, NON_UNIT, n,
                                                             L.win1(dx.data_ptr(), 1), L.win2(dA.data_ptr(), n))
    assert st == 0
    assert np.array_equal(dx.cpu().numpy(), x)


def _ratio(got, truth, mag, K, eps):
    return ((got.double() - truth).abs() / (mag * (K + 2) * eps).clamp_min(1e-300)).max().item()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_syrk_config3(so, dt):
    """syrk lower-N, N = 4096, K = 1024, alpha 1.5 (beta by trmscal 0.5)."""
    import torch
    from exoblas_amd import cblas
    N, K = 4096, 1024
    tdt = torch.float32 if dt == np.float32 else torch.float64
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    A = torch.empty((N, K), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    Cm = torch.empty((N, N), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
    C0 = Cm.clone()
    f = cblas.exo_ssyrk if dt == np.float32 else cblas.exo_dsyrk
    f(RM, LOWER, NO_TRANS, N, K, 1.5, A, K, 0.5, Cm, N)
    torch.cuda.synchronize()
    mask = torch.ones((N, N), dtype=torch.bool, device="cuda").tril()
    X = A.double()
    truth = torch.where(mask, 0.5 * C0.double() + 1.5 * (X @ X.T), C0.double())
    mag = torch.where(mask, 0.5 * C0.double().abs() + 1.5 * (X.abs() @ X.abs().T), C0.double().abs())
    assert _ratio(Cm, truth, mag, K, float(np.finfo(dt).eps)) <= 1.0
    assert torch.equal(Cm[~mask], C0[~mask])


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_symm_config3(so, dt):
    """symm left-lower, M = N = 4096, alpha 1.5, beta 0.5."""
    import torch
    from exoblas_amd import cblas
    from oracle import LEFT
    M = 4096
    tdt = torch.float32 if dt == np.float32 else torch.float64
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    S, B, Cm = (torch.empty((M, M), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g) for _ in range(3))
    C0 = Cm.clone()
    f = cblas.exo_ssymm if dt == np.float32 else cblas.exo_dsymm
    f(RM, LEFT, LOWER, M, M, 1.5, S, M, B, M, 0.5, Cm, M)
    torch.cuda.synchronize()
    T = S.double().tril()
    Sf = T + T.T - torch.diag(torch.diag(T))
    truth = 0.5 * C0.double() + 1.5 * (Sf @ B.double())
    mag = 0.5 * C0.double().abs() + 1.5 * (Sf.abs() @ B.double().abs())
    assert _ratio(Cm, truth, mag, M, float(np.finfo(dt).eps)) <= 1.0


def test_syrk_split_config4(so):
    """ssyrk lower-N N = 8192, K = 2048, plain and split (3, 4, 6 products)
    on both sets.  Asserted: every form within its bound on the normwise
    metric, the 6-product form at most plain FP32's max relative error on
    both sets; both metrics printed for each form."""
    import torch
    from exoblas_amd import lib as L
    N, K = 8192, 2048
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    stream = torch.cuda.current_stream().cuda_stream
    one = ctypes.byref(ctypes.c_float(1.0))
    per = {3: 2.0 ** -16, 4: 2.0 ** -16, 6: 2.0 ** -23}
    for scaled in (False, True):
        A = torch.empty((N, K), device="cuda").uniform_(-1, 1, generator=g)
        if scaled:
            A *= torch.exp2(torch.empty((N, 1), device="cuda").uniform_(-8, 8, generator=g))
        A64 = A.double()
        tri = torch.ones((N, N), dtype=torch.bool, device="cuda").tril_()
        truth = (A64 @ A64.T)[tri]
        mag = (A64.abs() @ A64.abs().T)[tri]
        del A64
        C = torch.zeros((N, N), device="cuda")
        wa, wc = L.win2(A.data_ptr(), K), L.win2(C.data_ptr(), N)
        errs = {}
        for form in ("plain", 3, 4, 6):
            C.zero_()
            if form == "plain":
                st = so.exoblas_ssyrk_rm_stride_1_async(stream, LOWER, NO_TRANS, N, K, one, wa, wa, wa, wa, wc)
            else:
                st = so.exoblas_ssyrk_split_async(stream, LOWER, NO_TRANS, N, K, one, wa, wa, wc, form)
            assert st == 0, L.last_error()
            torch.cuda.synchronize()
            assert torch.count_nonzero(C.triu(1)) == 0  # upper triangle untouched
            err = (C[tri].double() - truth).abs()
            errs[form] = ((err / mag).max().item(), (err / truth.abs().clamp_min(1e-300)).max().item())
            del err
        print(f"ssyrk 8192x2048 {'scaled' if scaled else 'uniform'}: (normwise, max rel) {errs}")
        for form in (3, 4, 6):
            assert errs[form][0] <= 4 * (K + 2) * (EPS32 + per[form]), (form, errs)
        assert errs["plain"][0] <= 4 * (K + 2) * EPS32, errs
        assert errs[6][1] <= errs["plain"][1], errs
        del A, C, truth, mag, tri
        torch.cuda.empty_cache()


@pytest.mark.parametrize("update", [16, 32])
def test_trsv_mixed_config4(so, update):
    """strsv lower-N-nonunit n = 8192, BF16 / FP32 off-diagonal updates, FP64
    diagonal blocks, one refinement step: within 4 eps_32 |x| of the float64
    solve, and no worse than the plain FP32 solve's max error."""
    import scipy.linalg
    import torch
    from exoblas_amd import lib as L
    n = 8192
    rng = np.random.default_rng(0)
    A = dd_matrix(rng, n, np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    exact = scipy.linalg.solve_triangular(np.tril(A).astype(np.float64), b.astype(np.float64), lower=True)
    dA = torch.from_numpy(A).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    dx = torch.from_numpy(b).cuda()
    st = so.exoblas_strsv_mixed_async(stream, LOWER, NO_TRANS, NON_UNIT, n, L.win1(dx.data_ptr(), 1),
                                      L.win2(dA.data_ptr(), n), update, 1)
    assert st == 0, L.last_error()
    x = dx.cpu().numpy().astype(np.float64)
    err = np.abs(x - exact)
    assert np.all(err <= 4 * EPS32 * np.abs(exact) + 1e-30), float(np.max(err / (EPS32 * np.abs(exact))))
    dx.copy_(torch.from_numpy(b))
    st = so.exoblas_strsv_rm_stride_1_async(stream, LOWER, NO_TRANS, NON_UNIT, n, L.win1(dx.data_ptr(), 1),
                                            L.win2(dA.data_ptr(), n))
    assert st == 0
    plain = np.abs(dx.cpu().numpy().astype(np.float64) - exact)
    assert err.max() <= plain.max(), (err.max(), plain.max())
