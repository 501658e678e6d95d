"""Level-1 parity: the HIP kernels behind the exo_* C-ABI against the oracle.

Sweeps follow the reference's correctness tests: N in {2, 321, 1000} x
increments {1, 2/3, negative} (test/level1/axpy/correctness.cpp:30-44,
dot/correctness.cpp:30-41, sdsdot/correctness.cpp:26), alpha in
{1.2, 0, 1} for axpy and {0, 1, 2} for scal (scal/correctness.cpp:29), rotm
flags {-2, -1, 0, 1} (rotm/correctness.cpp:35), plus n = 0 and n = 2^20.
Calls go through exoblas_amd.cblas, which mirrors the reference wrappers, so
both the _stride_1 and the _stride_any symbols are exercised, on host
buffers (the drop-in path).  The _async device-pointer path is covered at the
end.  copy and swap are compared bit for bit; the rest within the K*eps bound.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import Oracle
from util import assert_bound, bits_equal, elems, vec

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
SIZES = [(2, 1, 1), (321, 1, 1), (1000, 1, 1), (2, 2, 2), (321, 3, 3), (1000, 2, -3), (1000, -2, -4),
         (0, 1, 1), (1 << 20, 1, 1), (4099, 1, 1), (4099, -1, -1), (1 << 20, -1, -1), (1001, -1, 1)]


@pytest.fixture(scope="module")
def C():
    from exoblas_amd import cblas, lib
    lib.load()
    return cblas


def fn(C, dt, name):
    return getattr(C, ("exo_s" if dt == np.float32 else "exo_d") + name)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ix,iy", SIZES)
def test_copy_swap(C, dt, n, ix, iy, rng):
    x = vec(rng, n, ix, dt)
    y = vec(rng, n, iy, dt)
    yo = y.copy()
    Oracle(dt).copy(n, x, ix, yo, iy)
    fn(C, dt, "copy")(n, x, ix, y, iy)
    assert bits_equal(y, yo)
    x = vec(rng, n, ix, dt)
    y = vec(rng, n, iy, dt)
    xo, yo = x.copy(), y.copy()
    Oracle(dt).swap(n, xo, ix, yo, iy)
    fn(C, dt, "swap")(n, x, ix, y, iy)
    assert bits_equal(x, xo) and bits_equal(y, yo)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ix,iy", SIZES)
@pytest.mark.parametrize("alpha", [1.2, 0.0, 1.0, 0.75])
def test_axpy(C, dt, n, ix, iy, alpha, rng):
    x = vec(rng, n, ix, dt)
    y = vec(rng, n, iy, dt)
    yo = y.copy()
    if alpha != 0.0:  # the wrapper returns early for alpha == 0 (exo_axpy_wrapper.h:8-10)
        if alpha == 1.0 and ix == 1 and iy == 1:
            Oracle(dt).axpy_alpha_1(n, x, ix, yo, iy)
        else:
            Oracle(dt).axpy(n, alpha, x, ix, yo, iy)
    fn(C, dt, "axpy")(n, alpha, x, ix, y, iy)
    assert_bound(y, yo, np.abs(yo) + 2.0, 1, dt, what="axpy")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ix,iy", SIZES)
def test_dot_asum(C, dt, n, ix, iy, rng):
    x = vec(rng, n, ix, dt)
    y = vec(rng, n, iy, dt)
    o = Oracle(dt)
    xe, ye = elems(x, n, ix).astype(np.float64), elems(y, n, iy).astype(np.float64)
    got = fn(C, dt, "dot")(n, x, ix, y, iy)
    assert_bound(got, o.dot(n, x, ix, y, iy), np.sum(np.abs(xe * ye)), n, dt, what="dot")
    got = fn(C, dt, "asum")(n, x, ix)
    want = o.asum(n, x, ix) if ix > 0 else 0.0  # asum returns 0 for incX < 0 (exo_asum_wrapper.h)
    assert_bound(got, want, np.sum(np.abs(xe)), n, dt, what="asum")


@pytest.mark.parametrize("n,ix,iy", SIZES)
def test_dsdot_sdsdot(C, n, ix, iy, rng):
    x = vec(rng, n, ix, np.float32)
    y = vec(rng, n, iy, np.float32)
    o = Oracle(np.float32)
    xe, ye = elems(x, n, ix).astype(np.float64), elems(y, n, iy).astype(np.float64)
    mag = np.sum(np.abs(xe * ye))
    got = C.exo_dsdot(n, x, ix, y, iy)
    assert_bound(got, o.dsdot(n, x, ix, y, iy), mag, n, np.float64, what="dsdot")
    got = C.exo_sdsdot(n, 0.25, x, ix, y, iy)
    want = o.sdsdot(n, 0.25, x, ix, y, iy)
    # one rounding to fp32 after an fp64 sum: at most 1 ulp of fp32 apart
    assert abs(got - want) <= 2.0 ** -23 * (abs(want) + 1e-30) + n * 2.0 ** -52 * mag


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ix,iy", SIZES)
def test_rot(C, dt, n, ix, iy, rng):
    x = vec(rng, n, ix, dt)
    y = vec(rng, n, iy, dt)
    xo, yo = x.copy(), y.copy()
    Oracle(dt).rot(n, xo, ix, yo, iy, 0.6, 0.8)
    fn(C, dt, "rot")(n, x, ix, y, iy, 0.6, 0.8)
    assert_bound(x, xo, np.abs(xo) + 1.5, 2, dt, what="rot x")
    assert_bound(y, yo, np.abs(yo) + 1.5, 2, dt, what="rot y")


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ix,iy", SIZES[:8])
@pytest.mark.parametrize("flag", [-2.0, -1.0, 0.0, 1.0])
def test_rotm(C, dt, n, ix, iy, flag, rng):
    x = vec(rng, n, ix, dt)
    y = vec(rng, n, iy, dt)
    P = np.array([flag, 0.3, -0.7, 1.1, 0.45], dtype=dt)
    xo, yo = x.copy(), y.copy()
    if flag != -2.0:  # flag -2: no call (exo_rotm_wrapper.h:10-21)
        Oracle(dt).rotm(int(flag), n, xo, ix, yo, iy, np.array([P[1], P[3], P[2], P[4]], dtype=dt))
    fn(C, dt, "rotm")(n, x, ix, y, iy, P)
    assert_bound(x, xo, np.abs(xo) + 2.0, 2, dt, what="rotm x")
    assert_bound(y, yo, np.abs(yo) + 2.0, 2, dt, what="rotm y")
    if flag == -2.0:
        assert bits_equal(x, xo) and bits_equal(y, yo)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n,ix,iy", SIZES)
@pytest.mark.parametrize("alpha", [0.0, 1.0, 2.0])
def test_scal(C, dt, n, ix, iy, alpha, rng):
    x = vec(rng, n, ix, dt)
    xo = x.copy()
    if alpha != 1.0 and ix > 0:  # alpha == 1 and incX < 0 return early (exo_scal_wrapper.h)
        if alpha == 0.0:
            Oracle(dt).scal_alpha_0(n, xo, ix)
        else:
            Oracle(dt).scal(n, alpha, xo, ix)
    fn(C, dt, "scal")(n, alpha, x, ix)
    assert bits_equal(x, xo)  # one exact-or-correctly-rounded product per element


# ---------------------------------------------------------------- device path

@pytest.mark.parametrize("dt", DT)
def test_async_device_pointers(C, dt, rng):
    """The _async symbols on torch device tensors and torch's current stream,
    at the bandwidth-set size 2^26 reduced to 2^22 to keep the oracle fast."""
    import torch
    from exoblas_amd import lib as L
    so = L.load()
    p = "s" if dt == np.float32 else "d"
    T = ctypes.c_float if dt == np.float32 else ctypes.c_double
    n = 1 << 22
    x = vec(rng, n, 1, dt)
    y = vec(rng, n, 1, dt)
    dx = torch.from_numpy(x).cuda()
    dy = torch.from_numpy(y).cuda()
    dres = torch.zeros(1, dtype=dx.dtype, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    st = getattr(so, f"exoblas_{p}axpy_stride_1_async")(s, n, ctypes.byref(T(0.75)), L.win1(dx.data_ptr(), 1),
                                                        L.win1(dy.data_ptr(), 1))
    assert st == 0
    st = getattr(so, f"exoblas_{p}dot_stride_1_async")(s, n, L.win1(dx.data_ptr(), 1), L.win1(dy.data_ptr(), 1),
                                                       ctypes.c_void_p(dres.data_ptr()))
    assert st == 0
    torch.cuda.synchronize()
    yo = y.copy()
    Oracle(dt).axpy(n, 0.75, x, 1, yo, 1)
    assert_bound(dy.cpu().numpy(), yo, np.abs(yo) + 2, 1, dt, what="async axpy")
    want = Oracle(dt).dot(n, x, 1, yo, 1)
    assert_bound(dres.item(), want, np.sum(np.abs(x.astype(np.float64) * yo)), n, dt, what="async dot")
    # scopy / sswap at 2^22, bit-exact
    gy = dy.cpu().numpy()
    dz = torch.empty_like(dx)
    assert getattr(so, f"exoblas_{p}copy_stride_1_async")(s, n, L.win1(dx.data_ptr(), 1),
                                                          L.win1(dz.data_ptr(), 1)) == 0
    assert getattr(so, f"exoblas_{p}swap_stride_1_async")(s, n, L.win1(dz.data_ptr(), 1),
                                                          L.win1(dy.data_ptr(), 1)) == 0
    torch.cuda.synchronize()
    assert bits_equal(dy.cpu().numpy(), x) and bits_equal(dz.cpu().numpy(), gy)


def test_invalid_arguments_report_status(C):
    from exoblas_amd import lib as L
    so = L.load()
    x = np.zeros(4, np.float32)
    so.exo_sgemv_rm_stride_1(None, 115, 2, 2, ctypes.byref(ctypes.c_float(1)), ctypes.byref(ctypes.c_float(0)),
                             L.win2(x.ctypes.data, 2), L.win2(x.ctypes.data, 2), L.win1(x.ctypes.data, 1),
                             L.win1(x.ctypes.data, 1))
    assert L.last_status() == 1  # EXOBLAS_STATUS_INVALID_ENUM
    L.clear_error()
    so.exo_saxpy_stride_1(None, -1, ctypes.byref(ctypes.c_float(1)), L.win1(x.ctypes.data, 1),
                          L.win1(x.ctypes.data, 1))
    assert L.last_status() == 2
    L.clear_error()
