"""CBLAS-style entry points over the exo_* C-ABI.

Each function mirrors the argument handling of the reference's C++ wrapper
for that kernel (test/level{1,2,3}/<k>/exo_<k>_wrapper.h): the same early
returns, the same choice between the ``_stride_1`` and ``_stride_any``
symbol, the same re-basing of negative increments
(``X = X + (1 - N) * incX``), the same ColMajor rejection.  Buffers are numpy
arrays (host memory, the drop-in path) or torch tensors (device memory); the
pointer passed is the buffer's first element, as in CBLAS.
"""
from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from . import lib as L

ROW_MAJOR, COL_MAJOR = 101, 102


class UnsupportedParameterException(RuntimeError):
    """Raised where the reference wrappers throw (Order != RowMajor)."""


def _addr(buf: Any) -> int:
    if isinstance(buf, np.ndarray):
        return buf.ctypes.data
    if hasattr(buf, "data_ptr"):
        return buf.data_ptr()
    if isinstance(buf, int):
        return buf
    raise TypeError(f"unsupported buffer type {type(buf)}")


def _rebase(addr: int, n: int, inc: int, esize: int) -> int:
    """X + (1 - N) * incX for a negative increment (exo_axpy_wrapper.h:23-28)."""
    return addr + (1 - n) * inc * esize if inc < 0 else addr


def _call(name: str, *args) -> None:
    lib = L.load()
    lib.exoblas_clear_error()  # the status is sticky: check reports this call only
    getattr(lib, name)(None, *args)
    L.check()


def _make(p: str):
    T = ctypes.c_float if p == "s" else ctypes.c_double
    es = ctypes.sizeof(T)
    ns = {}

    def ref(v):
        return ctypes.byref(T(v))

    # ---------------------------------------------------------------- level 1
    def asum(N, X, incX):  # test/level1/asum/exo_asum_wrapper.h:6-23
        res = T(0)
        if incX == 1:
            _call(f"exo_{p}asum_stride_1", N, L.win1(_addr(X), incX), ctypes.byref(res))
            return res.value
        if incX < 0:
            return 0.0
        _call(f"exo_{p}asum_stride_any", N, L.win1(_addr(X), incX), ctypes.byref(res))
        return res.value

    def axpy(N, alpha, X, incX, Y, incY):  # test/level1/axpy/exo_axpy_wrapper.h:6-34
        if alpha == 0.0:
            return
        if incX == 1 and incY == 1:
            if alpha == 1.0:
                _call(f"exo_{p}axpy_alpha_1_stride_1", N, L.win1(_addr(X), 1), L.win1(_addr(Y), 1))
            else:
                _call(f"exo_{p}axpy_stride_1", N, ref(alpha), L.win1(_addr(X), 1), L.win1(_addr(Y), 1))
        else:
            x = _rebase(_addr(X), N, incX, es)
            y = _rebase(_addr(Y), N, incY, es)
            _call(f"exo_{p}axpy_stride_any", N, ref(alpha), L.win1(x, incX), L.win1(y, incY))

    def dot(N, X, incX, Y, incY):  # test/level1/dot/exo_dot_wrapper.h:6-27
        res = T(0)
        if incX == 1 and incY == 1:
            _call(f"exo_{p}dot_stride_1", N, L.win1(_addr(X), 1), L.win1(_addr(Y), 1), ctypes.byref(res))
        else:
            x = _rebase(_addr(X), N, incX, es)
            y = _rebase(_addr(Y), N, incY, es)
            _call(f"exo_{p}dot_stride_any", N, L.win1(x, incX), L.win1(y, incY), ctypes.byref(res))
        return res.value

    def _pair(kind):
        def f(N, X, incX, Y, incY):  # copy / swap wrappers, e.g. exo_copy_wrapper.h:6-23
            if incX == 1 and incY == 1:
                _call(f"exo_{p}{kind}_stride_1", N, L.win1(_addr(X), 1), L.win1(_addr(Y), 1))
            else:
                x = _rebase(_addr(X), N, incX, es)
                y = _rebase(_addr(Y), N, incY, es)
                _call(f"exo_{p}{kind}_stride_any", N, L.win1(x, incX), L.win1(y, incY))
        return f

    def rot(N, X, incX, Y, incY, c, s):  # test/level1/rot/exo_rot_wrapper.h:6-23
        if incX == 1 and incY == 1:
            _call(f"exo_{p}rot_stride_1", N, L.win1(_addr(X), 1), L.win1(_addr(Y), 1), ref(c), ref(s))
        else:
            x = _rebase(_addr(X), N, incX, es)
            y = _rebase(_addr(Y), N, incY, es)
            _call(f"exo_{p}rot_stride_any", N, L.win1(x, incX), L.win1(y, incY), ref(c), ref(s))

    def rotm(N, X, incX, Y, incY, P):  # test/level1/rotm/exo_rotm_wrapper.h:6-44
        H = (T * 4)(P[1], P[3], P[2], P[4])
        flag = {-1.0: "neg_one", 0.0: "zero", 1.0: "one"}.get(float(P[0]))
        if flag is None:  # flag -2 (or anything else): no call
            return
        if incX == 1 and incY == 1:
            _call(f"exo_{p}rotm_flag_{flag}_stride_1", N, L.win1(_addr(X), 1), L.win1(_addr(Y), 1), H)
        else:
            x = _rebase(_addr(X), N, incX, es)
            y = _rebase(_addr(Y), N, incY, es)
            _call(f"exo_{p}rotm_flag_{flag}_stride_any", N, L.win1(x, incX), L.win1(y, incY), H)

    def scal(N, alpha, X, incX):  # test/level1/scal/exo_scal_wrapper.h:8-35
        if alpha == 1.0:
            return
        if incX == 1:
            if alpha == 0.0:
                _call(f"exo_{p}scal_alpha_0_stride_1", N, L.win1(_addr(X), 1))
            else:
                _call(f"exo_{p}scal_stride_1", N, ref(alpha), L.win1(_addr(X), 1))
        else:
            if incX < 0:
                return
            if alpha == 0.0:
                _call(f"exo_{p}scal_alpha_0_stride_any", N, L.win1(_addr(X), incX))
            else:
                _call(f"exo_{p}scal_stride_any", N, ref(alpha), L.win1(_addr(X), incX))

    # ---------------------------------------------------------------- level 2
    def gemv(order, TransA, M, N, alpha, A, lda, X, incX, beta, Y, incY):
        # test/level2/gemv/exo_gemv_wrapper.h:8-39
        if order != ROW_MAJOR:
            raise UnsupportedParameterException("Unsupported Exception")
        m = M if TransA == L.NO_TRANS else N
        n = N if TransA == L.NO_TRANS else M
        a = L.win2(_addr(A), lda)
        if incX == 1 and incY == 1:
            _call(f"exo_{p}gemv_rm_stride_1", TransA, m, n, ref(alpha), ref(beta), a, a,
                  L.win1(_addr(X), 1), L.win1(_addr(Y), 1))
        else:
            x = _rebase(_addr(X), n, incX, es)
            y = _rebase(_addr(Y), m, incY, es)
            _call(f"exo_{p}gemv_rm_stride_any", TransA, m, n, ref(alpha), ref(beta), a, a,
                  L.win1(x, incX), L.win1(y, incY))

    def ger(order, M, N, alpha, X, incX, Y, incY, A, lda):  # test/level2/ger/exo_ger_wrapper.h:8-34
        if alpha == 0.0:
            return
        a = L.win2(_addr(A), lda)
        if incX == 1 and incY == 1:
            _call(f"exo_{p}ger_rm_stride_1", M, N, ref(alpha), L.win1(_addr(X), 1), L.win1(_addr(Y), 1), a)
        else:
            x = _rebase(_addr(X), M, incX, es)
            y = _rebase(_addr(Y), N, incY, es)
            _call(f"exo_{p}ger_rm_stride_any", M, N, ref(alpha), L.win1(x, incX), L.win1(y, incY), a)

    def symv(order, Uplo, N, alpha, A, lda, X, incX, beta, Y, incY):
        # test/level2/symv/exo_symv_wrapper.h:8-34
        if order != ROW_MAJOR:
            raise UnsupportedParameterException("Unsupported Exception")
        x = _rebase(_addr(X), N, incX, es)
        y = _rebase(_addr(Y), N, incY, es)
        v = "stride_1" if incX == 1 and incY == 1 else "stride_any"
        _call(f"exo_{p}symv_rm_{v}", Uplo, N, ref(alpha), L.win2(_addr(A), lda), L.win1(x, incX), ref(beta),
              L.win1(y, incY))

    def syr(order, Uplo, N, alpha, X, incX, A, lda):  # test/level2/syr/exo_syr_wrapper.h:6-32
        if order != ROW_MAJOR:
            raise UnsupportedParameterException("Unsupported Exception")
        if alpha == 0.0:
            return
        x = _rebase(_addr(X), N, incX, es) if incX != 1 else _addr(X)
        v = "stride_1" if incX == 1 else "stride_any"
        _call(f"exo_{p}syr_rm_{v}", Uplo, N, ref(alpha), L.win1(x, incX), L.win1(x, incX),
              L.win2(_addr(A), lda))

    def syr2(order, Uplo, N, alpha, X, incX, Y, incY, A, lda):
        # test/level2/syr2/exo_syr2_wrapper.h:6-36 (no Order check there)
        if alpha == 0.0:
            return
        x = _rebase(_addr(X), N, incX, es)
        y = _rebase(_addr(Y), N, incY, es)
        v = "stride_1" if incX == 1 and incY == 1 else "stride_any"
        _call(f"exo_{p}syr2_rm_{v}", Uplo, N, ref(alpha), L.win1(x, incX), L.win1(x, incX),
              L.win1(y, incY), L.win1(y, incY), L.win2(_addr(A), lda))

    def _tr(kind):
        def f(order, Uplo, TransA, Diag, N, A, lda, X, incX):
            # test/level2/trmv/exo_trmv_wrapper.h:8-29, trsv/exo_trsv_wrapper.h:8-29
            if order != ROW_MAJOR:
                raise UnsupportedParameterException("Unsupported Exception")
            x = _rebase(_addr(X), N, incX, es)
            v = "stride_1" if incX == 1 else "stride_any"
            _call(f"exo_{p}{kind}_rm_{v}", Uplo, TransA, Diag, N, L.win1(x, incX), L.win2(_addr(A), lda))
        return f

    # ---------------------------------------------------------------- banded level 2
    def gbmv(M, N, KL, KU, alpha, A, lda, X, incX, beta, Y, incY):
        # test/level2/sgbmv/exo_sgbmv.h:5-20 (no Order argument, no re-basing)
        v = "stride_1" if incX == 1 and incY == 1 else "stride_any"
        _call(f"exo_{p}gbmv_row_major_NonTrans_{v}", M, N, KL, KU, ref(alpha), ref(beta), L.win2(_addr(A), lda),
              L.win1(_addr(X), incX), L.win1(_addr(Y), incY))

    def sbmv(order, Uplo, N, K, alpha, A, lda, X, incX, beta, Y, incY):
        # test/level2/ssbmv/exo_ssbmv.h:7-53: beta through scal_y when != 1, then the product
        if order != ROW_MAJOR:
            raise UnsupportedParameterException("Unsupported Exception")
        x = _rebase(_addr(X), N, incX, es)
        y = _rebase(_addr(Y), N, incY, es)
        if beta != 1.0:
            v = "stride_1" if incY == 1 else "stride_any"
            _call(f"exo_{p}sbmv_scal_y_{v}", N, ref(beta), L.win1(y, incY))
        u = "Upper" if Uplo == L.UPPER else "Lower"
        v = "stride_1" if incX == 1 and incY == 1 else "stride_any"
        _call(f"exo_{p}sbmv_row_major_{u}_{v}", N, K, ref(alpha), L.win2(_addr(A), lda), L.win1(x, incX),
              L.win1(y, incY))

    def _tb(kind):
        def f(order, Uplo, TransA, Diag, N, K, A, lda, X, incX):
            # test/level2/stbmv/exo_stbmv.h:6-70, stbsv/exo_stbsv.h:6-56; tbsv's
            # Trans forms exist as _stride_any only, whatever incX
            if order != ROW_MAJOR:
                raise UnsupportedParameterException("Unsupported Exception")
            x = _rebase(_addr(X), N, incX, es)
            u = "Upper" if Uplo == L.UPPER else "Lower"
            t = "NonTrans" if TransA == L.NO_TRANS else "Trans"
            one = incX == 1 and not (kind == "tbsv" and t == "Trans")
            v = "stride_1" if one else "stride_any"
            _call(f"exo_{p}{kind}_row_major_{u}_{t}_{v}", N, K, L.win1(x, incX), L.win2(_addr(A), lda),
                  int(Diag == L.UNIT))
        return f

    # ---------------------------------------------------------------- level 3
    def gemm(Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
        # test/level3/gemm/exo_gemm_wrapper.h:10-28: C *= beta, then C += alpha op(A) op(B)
        if Order != ROW_MAJOR:
            raise UnsupportedParameterException("sgemm::Order must be Row Major")
        c = L.win2(_addr(C), ldc)
        _call(f"exo_{p}mscal_rm_stride_1", M, N, ref(beta), c)
        a, b = L.win2(_addr(A), lda), L.win2(_addr(B), ldb)
        _call(f"exo_{p}gemm_stride_1", TransA, TransB, M, N, K, ref(alpha), a, b, a, b, c)

    def symm(Order, Side, Uplo, M, N, alpha, A, lda, B, ldb, beta, C, ldc):
        # test/level3/symm/exo_symm_wrapper.h:10-27
        if Order != ROW_MAJOR:
            raise UnsupportedParameterException("symm::Order must be Row Major")
        c = L.win2(_addr(C), ldc)
        _call(f"exo_{p}mscal_rm_stride_1", M, N, ref(beta), c)
        a = L.win2(_addr(A), lda)
        _call(f"exo_{p}symm_rm_stride_1", Side, Uplo, M, N, ref(alpha), a, a, L.win2(_addr(B), ldb), c)

    def syrk(Order, Uplo, Trans, N, K, alpha, A, lda, beta, C, ldc):
        # test/level3/syrk/exo_syrk_wrapper.h:10-28: trmscal(beta) then syrk
        if Order != ROW_MAJOR:
            raise UnsupportedParameterException("syrk::Order must be Row Major")
        c = L.win2(_addr(C), ldc)
        _call(f"exo_{p}trmscal_rm_stride_1", Uplo, N, ref(beta), c)
        a = L.win2(_addr(A), lda)
        _call(f"exo_{p}syrk_rm_stride_1", Uplo, Trans, N, K, ref(alpha), a, a, a, a, c)

    def syr2k(Order, Uplo, Trans, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
        # test/level3/syr2k/exo_syr2k_wrapper.h:9-32: trmscal(beta) then syr2k
        if Order != ROW_MAJOR:
            raise UnsupportedParameterException("syr2k::Order must be Row Major")
        c = L.win2(_addr(C), ldc)
        _call(f"exo_{p}trmscal_rm_stride_1", Uplo, N, ref(beta), c)
        a, b = L.win2(_addr(A), lda), L.win2(_addr(B), ldb)
        _call(f"exo_{p}syr2k_rm_stride_1", Uplo, Trans, N, K, ref(alpha), a, a, b, b, a, a, b, b, c)

    def trmm(Order, Side, Uplo, TransA, Diag, M, N, alpha, A, lda, B, ldb):
        # test/level3/strmm/exo_strmm.h:7-56: Left side only (Right: no call),
        # Diag passed as (Diag == CblasUnit); Trans and ConjTrans both transpose
        if Order != ROW_MAJOR:
            raise UnsupportedParameterException("Unsupported Exception")
        if Side != L.LEFT:
            return
        u = "Upper" if Uplo == L.UPPER else "Lower"
        t = "NonTrans" if TransA == L.NO_TRANS else "Trans"
        _call(f"exo_{p}trmm_row_major_Left_{u}_{t}", M, N, ref(alpha), L.win2(_addr(A), lda),
              L.win2(_addr(B), ldb), int(Diag == L.UNIT))

    ns.update(asum=asum, axpy=axpy, dot=dot, copy=_pair("copy"), swap=_pair("swap"), rot=rot, rotm=rotm,
              scal=scal, gemv=gemv, ger=ger, symv=symv, syr=syr, syr2=syr2, trmv=_tr("trmv"),
              trsv=_tr("trsv"), gemm=gemm, symm=symm, syrk=syrk, syr2k=syr2k, gbmv=gbmv, sbmv=sbmv,
              tbmv=_tb("tbmv"), tbsv=_tb("tbsv"))
    if p == "s":  # the reference exports strmm only (src/level3/trmm.py:133-142)
        ns.update(trmm=trmm)
    return {f"exo_{p}{k}": v for k, v in ns.items()}


def exo_dsdot(N, X, incX, Y, incY):  # test/level1/dsdot/exo_dsdot_wrapper.h:5-25
    res = ctypes.c_double(0)
    if incX == 1 and incY == 1:
        _call("exo_dsdot_stride_1", N, L.win1(_addr(X), 1), L.win1(_addr(Y), 1), ctypes.byref(res))
    else:
        x = _rebase(_addr(X), N, incX, 4)
        y = _rebase(_addr(Y), N, incY, 4)
        _call("exo_dsdot_stride_any", N, L.win1(x, incX), L.win1(y, incY), ctypes.byref(res))
    return res.value


def exo_sdsdot(N, alpha, X, incX, Y, incY):  # test/level1/sdsdot/exo_sdsdot.h:5-26
    res = ctypes.c_float(0)
    sb = ctypes.c_float(alpha)
    if incX == 1 and incY == 1:
        _call("exo_sdsdot_stride_1", N, ctypes.byref(sb), L.win1(_addr(X), 1), L.win1(_addr(Y), 1),
              ctypes.byref(res))
    else:
        x = _rebase(_addr(X), N, incX, 4)
        y = _rebase(_addr(Y), N, incY, 4)
        _call("exo_sdsdot_stride_any", N, ctypes.byref(sb), L.win1(x, incX), L.win1(y, incY),
              ctypes.byref(res))
    return res.value


globals().update(_make("s"))
globals().update(_make("d"))
