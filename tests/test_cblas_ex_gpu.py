"""CBLAS-shaped extension entry points (exoblas_{s,d}gemm_ex, gemv_ex): row- and
column-major storage, every transpose pair, beta fused into the product, C / y
not read when beta == 0 (NaN there must not reach the result), alpha == 0 and
K == 0 leaving beta C.  Checked against a float64 product of the logical
matrices, built from the stored arrays independently of the library's own
mapping (column-major M x N = the transpose of a row-major N x ld array), to
the K * eps bound; padding columns must stay untouched.

The exo_* gemm path at shapes that take the narrow (32 x 32 tile, K shared
across waves) kernel is held to the oracle the same way as the reference's
sweep (test/level3/gemm/correctness.cpp:45-64).
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np
import pytest

from oracle import NO_TRANS, TRANS, Oracle
from util import assert_bound, bits_equal, uniform

pytestmark = pytest.mark.gpu

ROW, COL = 101, 102
DT = [np.float32, np.float64]


@pytest.fixture(scope="module")
def so():
    from exoblas_amd import lib as L
    return L.load()


def _p(dt):
    return "s" if dt == np.float32 else "d"


def _sc(dt, v):
    return ctypes.byref((ctypes.c_float if dt == np.float32 else ctypes.c_double)(v))


def _stored_shape(order, trans, r, c):
    """Rows x cols of the stored array of an r x c logical operand op(X) = X
    (NoTrans) or X^T (Trans), in CBLAS order."""
    plain = (order == ROW) == (trans == NO_TRANS)
    return (r, c) if plain else (c, r)


def _logical(S, order, trans, r, c):
    """op(X) as an r x c float64 matrix, from the stored array S."""
    rows, cols = _stored_shape(order, trans, r, c)
    L = S[:rows, :cols].astype(np.float64)  # the row-major view of the storage
    X = L if order == ROW else L.T            # CBLAS's X
    return X if trans == NO_TRANS else X.T


def _C_view(S, order, M, N):
    L = S[:, :(N if order == ROW else M)].astype(np.float64)
    return L if order == ROW else L.T


GEMM_DIMS = [(1, 7, 32), (7, 1, 257), (64, 32, 7), (257, 64, 33), (512, 512, 512), (500, 470, 333),
             (384, 512, 1024), (33, 1000, 300)]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("ta,tb", list(itertools.product([NO_TRANS, TRANS], [NO_TRANS, TRANS])))
@pytest.mark.parametrize("M,N,K", GEMM_DIMS)
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_gemm_ex(so, dt, order, ta, tb, M, N, K, beta, rng):
    from exoblas_amd import lib as L
    pad = 3
    ra, ca = _stored_shape(order, ta, M, K)
    rb, cb = _stored_shape(order, tb, K, N)
    rc, cc = (M, N) if order == ROW else (N, M)
    A = uniform(rng, (ra, ca + pad), dt)
    B = uniform(rng, (rb, cb + pad), dt)
    Cs = uniform(rng, (rc, cc + pad), dt)
    if beta == 0.0:
        Cs[:, :cc] = np.nan  # must not be read
    C0 = Cs.copy()
    alpha = 1.5
    st = getattr(so, f"exoblas_{_p(dt)}gemm_ex")(None, order, ta, tb, M, N, K, _sc(dt, alpha),
                                                 L.win2(A.ctypes.data, A.shape[1]), L.win2(B.ctypes.data, B.shape[1]),
                                                 _sc(dt, beta), L.win2(Cs.ctypes.data, Cs.shape[1]))
    assert st == 0, L.last_error()
    opA, opB = _logical(A, order, ta, M, K), _logical(B, order, tb, K, N)
    want = alpha * (opA @ opB)
    mag = abs(alpha) * (np.abs(opA) @ np.abs(opB))
    if beta != 0.0:
        Cl = _C_view(C0, order, M, N)
        want = want + beta * Cl
        mag = mag + abs(beta) * np.abs(Cl)
    got = _C_view(Cs, order, M, N)
    assert np.all(np.isfinite(got))
    assert_bound(got, want, mag, K, dt, what=f"gemm_ex order {order}")
    assert bits_equal(Cs[:, cc:], C0[:, cc:])  # padding untouched


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("alpha,K", [(0.0, 64), (1.5, 0)])
@pytest.mark.parametrize("beta", [0.0, 1.0, -2.0])
def test_gemm_ex_beta_only(so, dt, order, alpha, K, beta, rng):
    """alpha == 0 or K == 0: C = beta C (zeros for beta == 0, NaN included;
    untouched for beta == 1)."""
    from exoblas_amd import lib as L
    M, N, pad = 37, 70, 2
    A = uniform(rng, (max(M, K, 1), max(M, K, 1) + pad), dt)
    B = uniform(rng, (max(N, K, 1), max(N, K, 1) + pad), dt)
    rc, cc = (M, N) if order == ROW else (N, M)
    Cs = uniform(rng, (rc, cc + pad), dt)
    Cs[0, 0] = np.nan
    C0 = Cs.copy()
    st = getattr(so, f"exoblas_{_p(dt)}gemm_ex")(None, order, NO_TRANS, NO_TRANS, M, N, K, _sc(dt, alpha),
                                                 L.win2(A.ctypes.data, A.shape[1]), L.win2(B.ctypes.data, B.shape[1]),
                                                 _sc(dt, beta), L.win2(Cs.ctypes.data, Cs.shape[1]))
    assert st == 0, L.last_error()
    if beta == 0.0:
        assert np.all(Cs[:, :cc] == 0)
    elif beta == 1.0:
        assert bits_equal(Cs, C0)
    else:
        want = (beta * C0[:, :cc]).astype(dt)
        assert np.isnan(Cs[0, 0])
        Cs[0, 0] = want[0, 0] = 0
        assert bits_equal(Cs[:, :cc], want)
    assert bits_equal(Cs[:, cc:], C0[:, cc:])


@pytest.mark.parametrize("dt", DT)
def test_gemm_ex_device_single_launch_repeatable(so, dt):
    """The 512^3 beta product on device memory through the _async form: the
    same bits on every call (partial tiles summed in a fixed order)."""
    import torch
    from exoblas_amd import lib as L
    n = 512
    tdt = torch.float32 if dt == np.float32 else torch.float64
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    A, B, C = (torch.empty((n, n), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g) for _ in range(3))
    s = torch.cuda.current_stream().cuda_stream
    outs = []
    for _ in range(3):
        Cw = C.clone()
        st = getattr(so, f"exoblas_{_p(dt)}gemm_ex_async")(s, ROW, NO_TRANS, NO_TRANS, n, n, n, _sc(dt, 1.5),
                                                           L.win2(A.data_ptr(), n), L.win2(B.data_ptr(), n),
                                                           _sc(dt, 0.5), L.win2(Cw.data_ptr(), n))
        assert st == 0
        outs.append(Cw.cpu().numpy())
    assert bits_equal(outs[0], outs[1]) and bits_equal(outs[0], outs[2])
    Ad, Bd, Cd = (t.double().cpu().numpy() for t in (A, B, C))
    want = 1.5 * (Ad @ Bd) + 0.5 * Cd
    mag = 1.5 * (np.abs(Ad) @ np.abs(Bd)) + 0.5 * np.abs(Cd)
    assert_bound(outs[0], want, mag, n, dt, what="gemm_ex 512^3 device")


NARROW_DIMS = [(512, 512, 512), (500, 470, 333), (384, 512, 1024), (512, 256, 4096), (480, 520, 300)]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("ta,tb", list(itertools.product([NO_TRANS, TRANS], [NO_TRANS, TRANS])))
@pytest.mark.parametrize("M,N,K", NARROW_DIMS)
def test_exo_gemm_narrow_shapes(so, dt, ta, tb, M, N, K, rng):
    """The exo_* gemm (C += alpha op(A) op(B), no beta) at shapes the narrow
    kernel serves, against the oracle."""
    from exoblas_amd import cblas
    ld_diff = 5
    A = uniform(rng, (M if ta == NO_TRANS else K, (K if ta == NO_TRANS else M) + ld_diff), dt)
    B = uniform(rng, (K if tb == NO_TRANS else N, (N if tb == NO_TRANS else K) + ld_diff), dt)
    Cm = uniform(rng, (M, N + ld_diff), dt)
    Co = Cm.copy()
    o = Oracle(dt)
    o.mscal(M, N, 0.25, Co, Cm.shape[1])
    o.gemm(ta, tb, M, N, K, 1.5, A, A.shape[1], B, B.shape[1], Co, Cm.shape[1])
    C0 = Cm.copy()
    getattr(cblas, f"exo_{_p(dt)}gemm")(ROW, ta, tb, M, N, K, 1.5, A, A.shape[1], B, B.shape[1], 0.25, Cm,
                                       Cm.shape[1])
    opA = np.abs(A[:, :K] if ta == NO_TRANS else A[:, :M].T).astype(np.float64)
    opB = np.abs(B[:, :N] if tb == NO_TRANS else B[:, :K].T).astype(np.float64)
    mag = 0.25 * np.abs(C0[:, :N]) + 1.5 * (opA @ opB)
    assert_bound(Cm[:, :N], Co[:, :N], mag, K, dt, what="exo gemm narrow")
    assert bits_equal(Cm[:, N:], C0[:, N:])


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side", [141, 142])
@pytest.mark.parametrize("uplo", [121, 122])
def test_exo_symm_at_narrow_shape(so, dt, side, uplo, rng):
    """symm at a shape whose plain product would take the narrow kernel: its
    symmetric operand forms must keep the tiled path."""
    from exoblas_amd import cblas
    M, N = 512, 480
    na = M if side == 141 else N
    A = uniform(rng, (na, na + 3), dt)
    B = uniform(rng, (M, N + 3), dt)
    Cm = uniform(rng, (M, N + 3), dt)
    Co = Cm.copy()
    o = Oracle(dt)
    o.mscal(M, N, 0.5, Co, Cm.shape[1])
    o.symm(side, uplo, M, N, 1.5, A, A.shape[1], B, B.shape[1], Co, Cm.shape[1])
    C0 = Cm.copy()
    getattr(cblas, f"exo_{_p(dt)}symm")(ROW, side, uplo, M, N, 1.5, A, A.shape[1], B, B.shape[1], 0.5, Cm,
                                       Cm.shape[1])
    S = A[:, :na].astype(np.float64)
    S = np.tril(S) + np.tril(S, -1).T if uplo == 122 else np.triu(S) + np.triu(S, 1).T
    Bd = np.abs(B[:, :N].astype(np.float64))
    mag = 0.5 * np.abs(C0[:, :N]) + 1.5 * (np.abs(S) @ Bd if side == 141 else Bd @ np.abs(S))
    assert_bound(Cm[:, :N], Co[:, :N], mag, na, dt, what="exo symm 512")


def _vec(rng, n, inc, dt):
    size = 1 + (n - 1) * abs(inc) if n > 0 else 1
    return uniform(rng, size, dt)


def _win1_addr(buf, n, inc):
    """Window data pointer: element 0 of the logical vector (the last
    physical element for a negative increment)."""
    esz = buf.dtype.itemsize
    return buf.ctypes.data + ((n - 1) * abs(inc) * esz if inc < 0 and n > 0 else 0)


def _logical_vec(buf, n, inc):
    idx = np.arange(n) * abs(inc)
    return buf[idx[::-1]] if inc < 0 else buf[idx]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("order", [ROW, COL])
@pytest.mark.parametrize("ta", [NO_TRANS, TRANS])
@pytest.mark.parametrize("M,N", [(1, 1), (7, 33), (257, 64), (1000, 3000), (3000, 1000)])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -3), (-1, 2)])
@pytest.mark.parametrize("beta", [0.0, 0.75])
def test_gemv_ex(so, dt, order, ta, M, N, incx, incy, beta, rng):
    from exoblas_amd import lib as L
    pad = 3
    rs, cs = (M, N) if order == ROW else (N, M)
    A = uniform(rng, (rs, cs + pad), dt)
    nx, ny = (N, M) if ta == NO_TRANS else (M, N)
    x = _vec(rng, nx, incx, dt)
    y = _vec(rng, ny, incy, dt)
    if beta == 0.0:
        y[:] = np.nan
    y0 = y.copy()
    alpha = -1.25
    st = getattr(so, f"exoblas_{_p(dt)}gemv_ex")(None, order, ta, M, N, _sc(dt, alpha),
                                                 L.win2(A.ctypes.data, A.shape[1]),
                                                 L.win1(_win1_addr(x, nx, incx), incx), _sc(dt, beta),
                                                 L.win1(_win1_addr(y, ny, incy), incy))
    assert st == 0, L.last_error()
    Lm = A[:, :cs].astype(np.float64)
    Am = Lm if order == ROW else Lm.T  # the M x N matrix
    op = Am if ta == NO_TRANS else Am.T
    xv = _logical_vec(x, nx, incx).astype(np.float64)
    want = alpha * (op @ xv)
    mag = abs(alpha) * (np.abs(op) @ np.abs(xv))
    if beta != 0.0:
        yl = _logical_vec(y0, ny, incy).astype(np.float64)
        want = want + beta * yl
        mag = mag + abs(beta) * np.abs(yl)
    got = _logical_vec(y, ny, incy)
    assert np.all(np.isfinite(got))
    assert_bound(got, want, mag, nx, dt, what=f"gemv_ex order {order} trans {ta}")
    # elements between the strided ones untouched
    mask = np.ones(y.size, bool)
    mask[np.arange(ny) * abs(incy)] = False
    assert bits_equal(y[mask], y0[mask])
