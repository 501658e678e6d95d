"""Device windows that the host staging path never produces: unit-stride
windows off 16-byte alignment (the same offset in x and y, which the kernels
peel to reach the vector path, and different offsets, which go element by
element), and strided windows, through the _async entry points on device
memory.  Each result is held to the oracle run on the same elements: copy,
swap and scal bit for bit, the rest within the K*eps bound.  gemv with a
strided or misaligned x (gathered to contiguous scratch before the row
kernels) is covered the same way.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import NO_TRANS, TRANS, Oracle
from util import assert_bound, bits_equal

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
# (n, x offset, y offset, incx, incy): offsets in elements from a 256-byte aligned base
CASES = [(1, 1, 1, 1, 1), (3, 1, 1, 1, 1), (5, 3, 3, 1, 1), (1000, 1, 1, 1, 1), (4099, 2, 2, 1, 1),
         (4099, 1, 3, 1, 1), (1 << 20, 3, 3, 1, 1), (1 << 20, 1, 0, 1, 1), (1000, 0, 0, 2, -3),
         (4099, 1, 2, -2, 3), (1 << 20, 0, 1, 2, 2), (7, 1, 1, -1, -1)]


@pytest.fixture(scope="module")
def env():
    import torch
    from exoblas_amd import lib as L
    return torch, L.load(), L


def _p(dt):
    return "s" if dt == np.float32 else "d"


def _sc(dt, v):
    return ctypes.byref((ctypes.c_float if dt == np.float32 else ctypes.c_double)(v))


class Buf:
    """A device buffer holding an n-vector at increment inc from element
    offset `off`, with its host twin."""

    def __init__(self, torch, rng, n, off, inc, dt):
        self.n, self.off, self.inc, self.dt = n, off, inc, dt
        size = off + 1 + (n - 1) * abs(inc) + 8
        self.host = rng.uniform(-1, 1, size).astype(dt)
        self.dev = torch.from_numpy(self.host.copy()).cuda()
        assert self.dev.data_ptr() % 256 == 0

    def first(self):
        """Physical index of element 0 (the highest for a negative increment)."""
        return self.off + ((self.n - 1) * abs(self.inc) if self.inc < 0 else 0)

    def win(self, L):
        return L.win1(self.dev.data_ptr() + self.first() * self.host.itemsize, self.inc)

    def oracle_view(self, arr):
        """The buffer from `off` on, as the oracle's padded-vector layout."""
        return arr[self.off:]

    def got(self):
        return self.dev.cpu().numpy()


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ox,oy,ix,iy", CASES)
def test_elementwise_windows(env, dt, n, ox, oy, ix, iy, rng):
    torch, so, L = env
    s = torch.cuda.current_stream().cuda_stream
    p = _p(dt)
    o = Oracle(dt)

    def pair():
        return Buf(torch, rng, n, ox, ix, dt), Buf(torch, rng, n, oy, iy, dt)

    # axpy
    x, y = pair()
    want = y.host.copy()
    o.axpy(n, 0.75, x.oracle_view(x.host), ix, y.oracle_view(want), iy)
    assert getattr(so, f"exoblas_{p}axpy_stride_any_async")(s, n, _sc(dt, 0.75), x.win(L), y.win(L)) == 0
    assert_bound(y.got(), want, np.abs(want) + 2.0, 1, dt, what="axpy")
    # copy
    x, y = pair()
    want = y.host.copy()
    o.copy(n, x.oracle_view(x.host), ix, y.oracle_view(want), iy)
    assert getattr(so, f"exoblas_{p}copy_stride_any_async")(s, n, x.win(L), y.win(L)) == 0
    assert bits_equal(y.got(), want)
    # swap
    x, y = pair()
    wx, wy = x.host.copy(), y.host.copy()
    o.swap(n, x.oracle_view(wx), ix, y.oracle_view(wy), iy)
    assert getattr(so, f"exoblas_{p}swap_stride_any_async")(s, n, x.win(L), y.win(L)) == 0
    assert bits_equal(x.got(), wx) and bits_equal(y.got(), wy)
    # rot
    x, y = pair()
    wx, wy = x.host.copy(), y.host.copy()
    o.rot(n, x.oracle_view(wx), ix, y.oracle_view(wy), iy, 0.6, 0.8)
    assert getattr(so, f"exoblas_{p}rot_stride_any_async")(s, n, x.win(L), y.win(L), _sc(dt, 0.6),
                                                            _sc(dt, 0.8)) == 0
    assert_bound(x.got(), wx, np.abs(wx) + 2.0, 2, dt, what="rot x")
    assert_bound(y.got(), wy, np.abs(wy) + 2.0, 2, dt, what="rot y")
    # scal (x only)
    x, _ = pair()
    want = x.host.copy()
    o.scal(n, 0.5, x.oracle_view(want), ix)
    assert getattr(so, f"exoblas_{p}scal_stride_any_async")(s, n, _sc(dt, 0.5), x.win(L)) == 0
    assert bits_equal(x.got(), want)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ox,oy,ix,iy", CASES)
def test_reduction_windows(env, dt, n, ox, oy, ix, iy, rng):
    torch, so, L = env
    s = torch.cuda.current_stream().cuda_stream
    p = _p(dt)
    o = Oracle(dt)
    x, y = Buf(torch, rng, n, ox, ix, dt), Buf(torch, rng, n, oy, iy, dt)
    res = torch.zeros(1, dtype=torch.float32 if dt == np.float32 else torch.float64, device="cuda")
    rp = ctypes.c_void_p(res.data_ptr())
    xi = x.host[x.off + np.arange(n) * abs(ix)][:: -1 if ix < 0 else 1].astype(np.float64)
    yi = y.host[y.off + np.arange(n) * abs(iy)][:: -1 if iy < 0 else 1].astype(np.float64)
    assert getattr(so, f"exoblas_{p}dot_stride_any_async")(s, n, x.win(L), y.win(L), rp) == 0
    want = o.dot(n, x.oracle_view(x.host), ix, y.oracle_view(y.host), iy)
    assert_bound(res.item(), want, np.sum(np.abs(xi * yi)), n, dt, what="dot")
    if ix > 0:
        assert getattr(so, f"exoblas_{p}asum_stride_any_async")(s, n, x.win(L), rp) == 0
        want = o.asum(n, x.oracle_view(x.host), ix)
        assert_bound(res.item(), want, np.sum(np.abs(xi)), n, dt, what="asum")
    # the same bits on a second call (fixed summation order)
    assert getattr(so, f"exoblas_{p}dot_stride_any_async")(s, n, x.win(L), y.win(L), rp) == 0
    first = res.item()
    assert getattr(so, f"exoblas_{p}dot_stride_any_async")(s, n, x.win(L), y.win(L), rp) == 0
    assert res.item() == first


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
@pytest.mark.parametrize("M,N", [(7, 33), (4096, 1000), (1000, 4096), (4096, 4096)])
@pytest.mark.parametrize("ox,ix,iy", [(0, 2, -3), (1, 1, 1), (3, -1, 2)])
def test_gemv_windows(env, dt, trans, M, N, ox, ix, iy, rng):
    torch, so, L = env
    s = torch.cuda.current_stream().cuda_stream
    p = _p(dt)
    rows, cols = (M, N) if trans == NO_TRANS else (N, M)
    A = rng.uniform(-1, 1, (rows, cols)).astype(dt)
    dA = torch.from_numpy(A).cuda()
    x = Buf(torch, rng, N, ox, ix, dt)
    y = Buf(torch, rng, M, 1, iy, dt)
    want = y.host.copy()
    Oracle(dt).gemv(trans, M, N, 1.5, 0.5, A, cols, x.oracle_view(x.host), ix, y.oracle_view(want), iy)
    assert getattr(so, f"exoblas_{p}gemv_rm_stride_any_async")(s, trans, M, N, _sc(dt, 1.5), _sc(dt, 0.5),
                                                                L.win2(dA.data_ptr(), cols),
                                                                L.win2(dA.data_ptr(), cols), x.win(L), y.win(L)) == 0
    xi = x.host[x.off + np.arange(N) * abs(ix)][:: -1 if ix < 0 else 1].astype(np.float64)
    op = A.astype(np.float64) if trans == NO_TRANS else A.astype(np.float64).T
    mag_y = np.zeros_like(want, dtype=np.float64)
    idx = y.off + np.arange(M) * abs(iy)
    idx = idx[::-1] if iy < 0 else idx
    mag_y[idx] = 1.5 * (np.abs(op) @ np.abs(xi)) + 0.5 * np.abs(y.host[idx])
    assert_bound(y.got(), want, mag_y, N, dt, what="gemv windows")
