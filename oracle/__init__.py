"""ctypes binding of the CPU oracle (oracle/exo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's CPU-baseline leg.  The product library
(exoblas_amd) never imports this package.

Three builds of the same source (oracle/Makefile):
  liboracle.so          strict restatement, IEEE order of operations, no FMA
                        contraction: ``s*`` (fp32) and ``d*`` (fp64) routines
                        plus ``x*`` (fp64 storage, long-double arithmetic: the
                        truth oracle errors are stated against);
  liboracle_avx512.so   the same loop nests built with the reference's avx512
                        preset flags, timed as the CPU baseline;
  liboracle_avx2.so     the same for hosts without AVX-512;
  liboracle_omp_*.so    the baseline loop nests on every host thread (OpenMP
                        over the independent outer loops).

Arrays are numpy arrays; vectors are passed as (array, offset-of-element-0,
stride) so negative strides address downwards from element 0, as the
reference callee sees them after its wrapper re-bases the pointer.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_I = ctypes.c_int64
_P = ctypes.c_void_p

NO_TRANS, TRANS = 111, 112
UPPER, LOWER = 121, 122
NON_UNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142


def build() -> None:
    r = subprocess.run(["make", "-s", "-C", HERE, "-B", "-j3"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")


def _host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return " avx512f" in flags and " avx512vl" in flags
    except OSError:
        return False


_libs = {}


def load(kind: str = "strict") -> ctypes.CDLL:
    """kind: "strict" (checker), "fast" (CPU baseline, one thread: avx512 or
    avx2 build) or "omp" (the same baseline on every host thread)."""
    if kind in _libs:
        return _libs[kind]
    isa = "avx512" if _host_has_avx512() else "avx2"
    if kind == "strict":
        name = "liboracle.so"
    elif kind == "omp":
        name = f"liboracle_omp_{isa}.so"
    else:
        name = f"liboracle_{isa}.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    for p, T in (("s", ctypes.c_float), ("d", ctypes.c_double), ("x", ctypes.c_double)):
        def d(name, res, *args):
            f = getattr(lib, f"oracle_{p}{name}")
            f.restype = res
            f.argtypes = list(args)
        d("asum", T, _I, _P, _I)
        d("axpy", None, _I, T, _P, _I, _P, _I)
        d("axpy_alpha_1", None, _I, _P, _I, _P, _I)
        d("dot", T, _I, _P, _I, _P, _I)
        d("copy", None, _I, _P, _I, _P, _I)
        d("swap", None, _I, _P, _I, _P, _I)
        d("rot", None, _I, _P, _I, _P, _I, T, T)
        d("rotm", None, ctypes.c_int, _I, _P, _I, _P, _I, _P)
        d("scal", None, _I, T, _P, _I)
        d("scal_alpha_0", None, _I, _P, _I)
        d("gemv", None, _I, _I, _I, T, T, _P, _I, _P, _I, _P, _I)
        d("ger", None, _I, _I, T, _P, _I, _P, _I, _P, _I)
        d("symv", None, _I, _I, T, _P, _I, _P, _I, T, _P, _I)
        d("syr", None, _I, _I, T, _P, _I, _P, _I)
        d("syr2", None, _I, _I, T, _P, _I, _P, _I, _P, _I)
        d("trmv", None, _I, _I, _I, _I, _P, _I, _P, _I)
        d("trsv", None, _I, _I, _I, _I, _P, _I, _P, _I)
        d("gemm", None, _I, _I, _I, _I, _I, T, _P, _I, _P, _I, _P, _I)
        d("mscal", None, _I, _I, T, _P, _I)
        d("trmscal", None, _I, _I, T, _P, _I)
        d("symm", None, _I, _I, _I, _I, T, _P, _I, _P, _I, _P, _I)
        d("syrk", None, _I, _I, _I, _I, T, _P, _I, _P, _I)
        d("syr2k", None, _I, _I, _I, _I, T, _P, _I, _P, _I, _P, _I)
        d("gbmv", None, _I, _I, _I, _I, T, T, _P, _I, _P, _I, _P, _I)
        d("sbmv_scal_y", None, _I, T, _P, _I)
        d("sbmv", None, _I, _I, _I, T, _P, _I, _P, _I, _P, _I)
        d("tbmv", None, _I, _I, _I, _I, _P, _I, _P, _I, _I)
        d("tbsv", None, _I, _I, _I, _I, _P, _I, _P, _I, _I)
        d("trmm", None, _I, _I, _I, _I, T, _P, _I, _P, _I, _I)
    lib.oracle_dsdot.restype = ctypes.c_double
    lib.oracle_dsdot.argtypes = [_I, _P, _I, _P, _I]
    lib.oracle_xdsdot.restype = ctypes.c_double
    lib.oracle_xdsdot.argtypes = [_I, _P, _I, _P, _I]
    lib.oracle_sdsdot.restype = ctypes.c_float
    lib.oracle_sdsdot.argtypes = [_I, ctypes.c_float, _P, _I, _P, _I]
    _libs[kind] = lib
    return lib


def prefix(dtype, truth: bool = False) -> str:
    if truth:
        return "x"
    return "s" if np.dtype(dtype) == np.float32 else "d"


def vptr(a: np.ndarray, n: int, inc: int) -> int:
    """Address of element 0 of a vector of n elements, stride inc, stored in a
    (the lowest-addressed element for inc < 0 is a[0]).  Raises if a is too
    short to hold the vector: the C restatement would write past its end."""
    need = (n - 1) * abs(inc) + 1 if n > 0 else 0
    if a.size < need:
        raise ValueError(f"vector of {n} at stride {inc} needs {need} elements, buffer has {a.size}")
    base = a.ctypes.data
    if inc < 0 and n > 0:
        base += (n - 1) * (-inc) * a.itemsize
    return base


def mptr(a: np.ndarray, rows: int, cols: int, ld: int) -> int:
    """Address of a row-major rows x cols matrix with leading dimension ld
    stored in a; raises if a is too short to hold it."""
    need = (rows - 1) * ld + cols if rows > 0 and cols > 0 else 0
    if a.size < need:
        raise ValueError(f"{rows} x {cols} matrix with ld {ld} needs {need} elements, buffer has {a.size}")
    return a.ctypes.data


class Oracle:
    """Calls the restated spec with numpy arrays, in place, like the callee.

    ``truth=True`` evaluates in long double on float64 copies; results are
    returned / written as float64 (the caller passes float64 arrays)."""

    def __init__(self, dtype=np.float32, truth: bool = False, kind: str = "strict"):
        self.lib = load(kind)
        self.p = prefix(dtype, truth)
        self.dtype = np.float64 if truth else np.dtype(dtype)

    def f(self, name):
        return getattr(self.lib, f"oracle_{self.p}{name}")

    def _chk(self, *arrs):
        for a in arrs:
            if a is not None:
                assert a.dtype == self.dtype and a.flags.c_contiguous, (a.dtype, self.dtype)

    # level 1 -----------------------------------------------------------
    def asum(self, n, x, incx=1):
        self._chk(x)
        return self.f("asum")(n, vptr(x, n, incx), incx)

    def axpy(self, n, alpha, x, incx, y, incy):
        self._chk(x, y)
        self.f("axpy")(n, alpha, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    def axpy_alpha_1(self, n, x, incx, y, incy):
        self._chk(x, y)
        self.f("axpy_alpha_1")(n, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    def dot(self, n, x, incx, y, incy):
        self._chk(x, y)
        return self.f("dot")(n, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    def copy(self, n, x, incx, y, incy):
        self.f("copy")(n, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    def swap(self, n, x, incx, y, incy):
        self.f("swap")(n, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    def rot(self, n, x, incx, y, incy, c, s):
        self._chk(x, y)
        self.f("rot")(n, vptr(x, n, incx), incx, vptr(y, n, incy), incy, c, s)

    def rotm(self, flag, n, x, incx, y, incy, H):
        self._chk(x, y)
        h = np.ascontiguousarray(H, dtype=self.dtype)
        self.f("rotm")(int(flag), n, vptr(x, n, incx), incx, vptr(y, n, incy), incy, h.ctypes.data)

    def scal(self, n, alpha, x, incx):
        self._chk(x)
        self.f("scal")(n, alpha, vptr(x, n, incx), incx)

    def scal_alpha_0(self, n, x, incx):
        self.f("scal_alpha_0")(n, vptr(x, n, incx), incx)

    def dsdot(self, n, x, incx, y, incy, truth=False):
        f = self.lib.oracle_xdsdot if truth else self.lib.oracle_dsdot
        return f(n, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    def sdsdot(self, n, sb, x, incx, y, incy):
        return self.lib.oracle_sdsdot(n, sb, vptr(x, n, incx), incx, vptr(y, n, incy), incy)

    # level 2 -----------------------------------------------------------
    def gemv(self, trans, m, n, alpha, beta, A, lda, x, incx, y, incy):
        """Callee view: m, n after the wrapper's swap (A is n x m for Trans)."""
        self._chk(A, x, y)
        rows, cols = (m, n) if trans == NO_TRANS else (n, m)
        self.f("gemv")(trans, m, n, alpha, beta, mptr(A, rows, cols, lda), lda, vptr(x, n, incx), incx,
                       vptr(y, m, incy), incy)

    def ger(self, m, n, alpha, x, incx, y, incy, A, lda):
        self._chk(A, x, y)
        self.f("ger")(m, n, alpha, vptr(x, m, incx), incx, vptr(y, n, incy), incy, mptr(A, m, n, lda), lda)

    def symv(self, uplo, n, alpha, A, lda, x, incx, beta, y, incy):
        self._chk(A, x, y)
        self.f("symv")(uplo, n, alpha, mptr(A, n, n, lda), lda, vptr(x, n, incx), incx, beta, vptr(y, n, incy),
                       incy)

    def syr(self, uplo, n, alpha, x, incx, A, lda):
        self._chk(A, x)
        self.f("syr")(uplo, n, alpha, vptr(x, n, incx), incx, mptr(A, n, n, lda), lda)

    def syr2(self, uplo, n, alpha, x, incx, y, incy, A, lda):
        self._chk(A, x, y)
        self.f("syr2")(uplo, n, alpha, vptr(x, n, incx), incx, vptr(y, n, incy), incy, mptr(A, n, n, lda), lda)

    def trmv(self, uplo, trans, diag, n, x, incx, A, lda):
        self._chk(A, x)
        self.f("trmv")(uplo, trans, diag, n, vptr(x, n, incx), incx, mptr(A, n, n, lda), lda)

    def trsv(self, uplo, trans, diag, n, x, incx, A, lda):
        self._chk(A, x)
        self.f("trsv")(uplo, trans, diag, n, vptr(x, n, incx), incx, mptr(A, n, n, lda), lda)

    # banded level 2 ----------------------------------------------------
    def gbmv(self, m, n, kl, ku, alpha, beta, A, lda, x, incx, y, incy):
        self._chk(A, x, y)
        self.f("gbmv")(m, n, kl, ku, alpha, beta, mptr(A, m, kl + ku + 1, lda), lda, vptr(x, n, incx), incx,
                       vptr(y, m, incy), incy)

    def sbmv_scal_y(self, n, beta, y, incy):
        self._chk(y)
        self.f("sbmv_scal_y")(n, beta, vptr(y, n, incy), incy)

    def sbmv(self, uplo, n, k, alpha, A, lda, x, incx, y, incy):
        """y += alpha S x (the proc after the wrapper's separate beta scaling)."""
        self._chk(A, x, y)
        self.f("sbmv")(int(uplo == UPPER), n, k, alpha, mptr(A, n, k + 1, lda), lda, vptr(x, n, incx), incx,
                       vptr(y, n, incy), incy)

    def tbmv(self, uplo, trans, diag, n, k, x, incx, A, lda):
        self._chk(A, x)
        self.f("tbmv")(int(uplo == UPPER), int(trans != NO_TRANS), n, k, vptr(x, n, incx), incx,
                       mptr(A, n, k + 1, lda), lda, int(diag == UNIT))

    def tbsv(self, uplo, trans, diag, n, k, x, incx, A, lda):
        self._chk(A, x)
        self.f("tbsv")(int(uplo == UPPER), int(trans != NO_TRANS), n, k, vptr(x, n, incx), incx,
                       mptr(A, n, k + 1, lda), lda, int(diag == UNIT))

    # level 3 -----------------------------------------------------------
    def gemm(self, ta, tb, M, N, K, alpha, A, lda, B, ldb, C, ldc):
        self._chk(A, B, C)
        ra, ca = (M, K) if ta == NO_TRANS else (K, M)
        rb, cb = (K, N) if tb == NO_TRANS else (N, K)
        self.f("gemm")(ta, tb, M, N, K, alpha, mptr(A, ra, ca, lda), lda, mptr(B, rb, cb, ldb), ldb,
                       mptr(C, M, N, ldc), ldc)

    def mscal(self, M, N, alpha, A, lda):
        self._chk(A)
        self.f("mscal")(M, N, alpha, mptr(A, M, N, lda), lda)

    def trmscal(self, uplo, N, alpha, A, lda):
        self._chk(A)
        self.f("trmscal")(uplo, N, alpha, mptr(A, N, N, lda), lda)

    def symm(self, side, uplo, M, N, alpha, A, lda, B, ldb, C, ldc):
        self._chk(A, B, C)
        na = M if side == LEFT else N
        self.f("symm")(side, uplo, M, N, alpha, mptr(A, na, na, lda), lda, mptr(B, M, N, ldb), ldb,
                       mptr(C, M, N, ldc), ldc)

    def syrk(self, uplo, trans, N, K, alpha, A, lda, C, ldc):
        self._chk(A, C)
        ra, ca = (N, K) if trans == NO_TRANS else (K, N)
        self.f("syrk")(uplo, trans, N, K, alpha, mptr(A, ra, ca, lda), lda, mptr(C, N, N, ldc), ldc)

    def syr2k(self, uplo, trans, N, K, alpha, A, lda, B, ldb, C, ldc):
        self._chk(A, B, C)
        ra, ca = (N, K) if trans == NO_TRANS else (K, N)
        self.f("syr2k")(uplo, trans, N, K, alpha, mptr(A, ra, ca, lda), lda, mptr(B, ra, ca, ldb), ldb,
                        mptr(C, N, N, ldc), ldc)

    def trmm(self, uplo, trans, diag, m, n, alpha, A, lda, B, ldb):
        """Left side: B <- alpha op(A) B (trmm.py's four Left procs); diag is
        the CBLAS value (UNIT reads no diagonal)."""
        self._chk(A, B)
        self.f("trmm")(int(uplo == UPPER), int(trans != NO_TRANS), m, n, alpha, mptr(A, m, m, lda), lda,
                       mptr(B, m, n, ldb), ldb, int(diag == UNIT))


def strided(n: int, inc: int, rng: np.random.Generator, dtype, fill: Optional[float] = None) -> np.ndarray:
    """A buffer holding an n-element vector at stride |inc| (padding included),
    filled with U[-1,1] (or `fill`), like the reference's AlignedBuffer
    (test/common/generate_buffer.h:12-101) but seeded."""
    size = max(1, 1 + (n - 1) * abs(inc)) if n > 0 else 1
    if fill is None:
        return rng.uniform(-1.0, 1.0, size).astype(dtype)
    return np.full(size, fill, dtype=dtype)
