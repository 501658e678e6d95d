"""Prototypes of libexoblas_amd.so's native entry points beyond the exo_* ABI
(the split-precision, mixed-precision, batched and serving APIs).

Each entry is (symbol, restype, argtypes).  ``declare`` sets them on the
loaded library; a symbol missing from the library is an error, so the Python
side and the C side cannot drift apart silently.
"""
from __future__ import annotations

import ctypes

_I = ctypes.c_int64
_P = ctypes.c_void_p
_i = ctypes.c_int

SYMBOLS = []


def declare(lib: ctypes.CDLL) -> None:
    for name, res, args in SYMBOLS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = list(args)
