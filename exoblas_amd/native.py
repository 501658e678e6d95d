"""Prototypes of libexoblas_amd.so's native entry points beyond the exo_* ABI
(include/exoblas_amd_batch.h: batched execution, plans, tuning table and the
serving loop), and the matching ctypes / numpy types.

Each SYMBOLS entry is (symbol, restype, argtypes).  ``declare`` sets them on
the loaded library; a symbol missing from the library is an error, so the
Python side and the C side cannot drift apart silently.
"""
from __future__ import annotations

import ctypes

import numpy as np

_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_P = ctypes.c_void_p
_i = ctypes.c_int

OP_GEMM, OP_GEMV, OP_TRSV = 1, 2, 3


class BatchOp(ctypes.Structure):
    """struct exoblas_batch_op"""
    _fields_ = [("kind", _I32), ("trans_a", _I32), ("trans_b", _I32), ("diag", _I32),
                ("m", _I64), ("n", _I64), ("k", _I64),
                ("a", _P), ("lda", _I64), ("b", _P), ("ldb", _I64), ("c", _P), ("ldc", _I64),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float)]


# The same layout as a numpy structured dtype, for building large op arrays.
OP_DTYPE = np.dtype({"names": [f[0] for f in BatchOp._fields_],
                     "formats": ["<i4"] * 4 + ["<i8"] * 3 + ["<u8", "<i8", "<u8", "<i8", "<u8", "<i8", "<f4", "<f4"],
                     "offsets": [getattr(BatchOp, f[0]).offset for f in BatchOp._fields_],
                     "itemsize": ctypes.sizeof(BatchOp)})


class ServerStats(ctypes.Structure):
    """struct exoblas_server_stats"""
    _fields_ = [("submitted", _I64), ("completed", _I64), ("rejected", _I64), ("batches", _I64),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double), ("mean_us", ctypes.c_double),
                ("max_us", ctypes.c_double), ("wall_s", ctypes.c_double), ("peak_bytes_in_flight", _I64),
                ("cancelled", _I64), ("failed", _I64)]


SYMBOLS = [
    ("exoblas_batch_run", _i, (_P, _P, _I64)),
    ("exoblas_batch_queue", _i, (_P, _I64, _P, _P, _I64, _P, _P)),
    ("exoblas_batch_plan_create", _i, (ctypes.POINTER(_P), _P, _I64)),
    ("exoblas_batch_plan_launch", _i, (_P, _P)),
    ("exoblas_batch_plan_destroy", _i, (_P,)),
    ("exoblas_batch_set_tuning", _i, (_P, _I64)),
    ("exoblas_batch_get_tuning", _i, (_P, _I64)),
    ("exoblas_batch_force_tile", _i, (_I32,)),
    ("exoblas_server_create", _i, (ctypes.POINTER(_P), _I64, _I32, _I32, _I32)),
    ("exoblas_server_submit", _I64, (_P, _P)),
    ("exoblas_server_replay", _i, (_P, _P, _P, _I64)),
    ("exoblas_server_drain", _i, (_P,)),
    ("exoblas_server_done", _i, (_P, _I64)),
    ("exoblas_server_cancel", _i, (_P, _I64)),
    ("exoblas_server_stats_get", _i, (_P, ctypes.POINTER(ServerStats))),
    ("exoblas_server_latencies", _I64, (_P, _P, _I64)),
    ("exoblas_server_destroy", _i, (_P,)),
]


def declare(lib: ctypes.CDLL) -> None:
    for name, res, args in SYMBOLS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = list(args)
