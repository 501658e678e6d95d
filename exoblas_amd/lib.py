"""ctypes binding of libexoblas_amd.so: the exo_* C-ABI as Python callables.

The argument types come from the same table that generates the C headers
(:mod:`exoblas_amd.generator`), so this binding and ``include/exo_*.h`` cannot
drift apart.  Loading fails loudly when the shared library is missing: there
is no CPU fallback behind these entry points.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

from . import generator

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexoblas_amd.so")

c_int_fast32 = ctypes.c_int64  # int_fast32_t is 8 bytes on x86-64 glibc

NO_TRANS, TRANS, CONJ_TRANS = 111, 112, 113
UPPER, LOWER = 121, 122
NON_UNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142


class Win1(ctypes.Structure):
    """struct exo_win_1{f32,f64}[c]: {T *data; int_fast32_t strides[1];}"""
    _fields_ = [("data", ctypes.c_void_p), ("strides", c_int_fast32 * 1)]


class Win2(ctypes.Structure):
    """struct exo_win_2{f32,f64}[c]: {T *data; int_fast32_t strides[2];}"""
    _fields_ = [("data", ctypes.c_void_p), ("strides", c_int_fast32 * 2)]


def win1(ptr: int, stride: int) -> Win1:
    return Win1(ptr, (c_int_fast32 * 1)(stride))


def win2(ptr: int, ld: int) -> Win2:
    return Win2(ptr, (c_int_fast32 * 2)(ld, 1))


class LibraryMissing(RuntimeError):
    pass


_lib: Optional[ctypes.CDLL] = None


def _ctype_of(p: str, a: generator.Arg):
    if a.kind in ("size", "enum"):
        return c_int_fast32
    if a.kind == "win":
        return Win1 if len(a.dims) == 1 else Win2
    return ctypes.c_void_p  # scalar / result / array pointers


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the library and declares every exported symbol's prototype."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(
            f"{path} is not built; run `python -m exoblas_amd.build` (the exo_* entry "
            "points have no CPU fallback)")
    # torch wheels bundle their own HIP runtime (libamdhip64.so.7,
    # libhsa-runtime64.so.1).  Loading torch first makes this library bind to
    # that same runtime by SONAME, so torch tensors and streams are valid here;
    # loaded the other way round the process would hold two HIP runtimes and
    # the second one finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for proc, p, v in generator.variants():
        argt = [_ctype_of(p, a) for a in proc.args]
        f = getattr(lib, generator.symbol(proc, p, v))
        f.argtypes = [ctypes.c_void_p] + argt
        f.restype = None
        g = getattr(lib, generator.async_symbol(proc, p, v))
        g.argtypes = [ctypes.c_void_p] + argt
        g.restype = ctypes.c_int
    for proc, p, v in generator.ext_variants():
        argt = [_ctype_of(p, a) for a in proc.args]
        for name in (generator.symbol(proc, p, v), generator.async_symbol(proc, p, v)):
            f = getattr(lib, name)
            f.argtypes = [ctypes.c_void_p] + argt
            f.restype = ctypes.c_int
    lib.exoblas_last_status.restype = ctypes.c_int
    lib.exoblas_last_error.restype = ctypes.c_char_p
    lib.exoblas_clear_error.restype = None
    lib.exoblas_init.restype = ctypes.c_int
    lib.exoblas_synchronize.restype = ctypes.c_int
    lib.exoblas_stream_check.restype = ctypes.c_int
    lib.exoblas_stream_check.argtypes = [ctypes.c_void_p]
    lib.exoblas_stream_release.restype = ctypes.c_int
    lib.exoblas_stream_release.argtypes = [ctypes.c_void_p]
    lib.exoblas_set_spin_limit.restype = ctypes.c_int
    lib.exoblas_set_spin_limit.argtypes = [ctypes.c_longlong]
    lib.exoblas_sdot_noalloc_async.restype = ctypes.c_int
    lib.exoblas_sdot_noalloc_async.argtypes = [ctypes.c_void_p, c_int_fast32, Win1, Win1, ctypes.c_void_p]
    lib.exoblas_strsv_mixed_pack_async.restype = ctypes.c_int
    lib.exoblas_strsv_mixed_pack_async.argtypes = [ctypes.c_void_p, c_int_fast32, Win2, ctypes.c_void_p]
    lib.exoblas_strsv_mixed_packed_async.restype = ctypes.c_int
    lib.exoblas_strsv_mixed_packed_async.argtypes = [ctypes.c_void_p] + [c_int_fast32] * 4 + [Win1, Win2, ctypes.c_void_p,
                                                                                             c_int_fast32]
    _extend(lib)
    _lib = lib
    return lib


def _extend(lib: ctypes.CDLL) -> None:
    """Prototypes of the library's native (non exo_*) entry points."""
    from . import native
    native.declare(lib)


def last_status() -> int:
    return load().exoblas_last_status()


def last_error() -> str:
    m = load().exoblas_last_error()
    return m.decode() if m else ""


def clear_error() -> None:
    load().exoblas_clear_error()


def check() -> None:
    """Raises if the last call on this thread failed."""
    st = last_status()
    if st != 0:
        msg = last_error()
        clear_error()
        raise RuntimeError(f"exoblas_amd status {st}: {msg}")


def symbols() -> Dict[str, ctypes._CFuncPtr]:
    lib = load()
    return {s: getattr(lib, s) for s in generator.all_symbols()}
