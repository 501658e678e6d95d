"""Host-buffer drop-in calls large enough to take the pinned, chunked
staging path (ranges of 4 MiB and more, 8 MiB chunks): results must equal
the device-path / oracle results, copies and swaps bit-exactly, across
uneven chunk boundaries and negative increments."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import NO_TRANS, TRANS, Oracle
from util import assert_bound

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    from exoblas_amd import cblas, lib
    lib.load()
    return cblas


@pytest.mark.parametrize("n,incx,incy", [((3 << 20) + 17, 1, 1), ((5 << 20) + 3, 1, -1), ((2 << 20) + 5, 2, 1)])
def test_copy_swap_large_host(C, n, incx, incy, rng):
    x = rng.uniform(-1, 1, n * abs(incx)).astype(np.float32)
    y = rng.uniform(-1, 1, n * abs(incy)).astype(np.float32)
    x0, y0 = x.copy(), y.copy()
    C.exo_sswap(n, x, incx, y, incy)
    xo, yo = x0.copy(), y0.copy()
    Oracle(np.float32).swap(n, xo, incx, yo, incy)
    assert np.array_equal(x.view(np.uint32), xo.view(np.uint32))
    assert np.array_equal(y.view(np.uint32), yo.view(np.uint32))
    z = np.zeros_like(y0)
    C.exo_scopy(n, x0, incx, z, incy)
    zo = np.zeros_like(y0)
    Oracle(np.float32).copy(n, x0, incx, zo, incy)
    assert np.array_equal(z.view(np.uint32), zo.view(np.uint32))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_axpy_dot_large_host(C, dt, rng):
    n = (3 << 20) + 1001
    x = rng.uniform(-1, 1, n).astype(dt)
    y = rng.uniform(-1, 1, n).astype(dt)
    yo = y.copy()
    p = "s" if dt == np.float32 else "d"
    getattr(C, f"exo_{p}axpy")(n, 0.75, x, 1, y, 1)
    Oracle(dt).axpy(n, 0.75, x, 1, yo, 1)
    assert_bound(y, yo, 0.75 * np.abs(x) + np.abs(yo), 1, dt, what="axpy")
    got = getattr(C, f"exo_{p}dot")(n, x, 1, y, 1)
    want = Oracle(dt).dot(n, x, 1, y, 1)
    mag = float(np.abs(x.astype(np.float64) * y).sum())
    assert abs(float(got) - float(want)) <= 4 * (n + 2) * float(np.finfo(dt).eps) * mag


@pytest.mark.parametrize("trans", [NO_TRANS, TRANS])
def test_gemv_large_host(C, trans, rng):
    M, N = 3001, 1503
    A = rng.uniform(-1, 1, (M, N + 3)).astype(np.float32)
    nx, ny = (N, M) if trans == NO_TRANS else (M, N)
    x = rng.uniform(-1, 1, nx).astype(np.float32)
    y = rng.uniform(-1, 1, ny).astype(np.float32)
    yo = y.copy()
    # the oracle takes the callee's view: m = len(y), n = len(x)
    Oracle(np.float32).gemv(trans, ny, nx, 1.5, 0.5, A, N + 3, x, 1, yo, 1)
    C.exo_sgemv(101, trans, M, N, 1.5, A, N + 3, x, 1, 0.5, y, 1)
    Ab = np.abs(A[:, :N].astype(np.float64))
    mag = 0.5 * np.abs(yo) + 1.5 * ((Ab @ np.abs(x)) if trans == NO_TRANS else (Ab.T @ np.abs(x)))
    assert_bound(y, yo, mag, nx, np.float32, what="gemv host")
