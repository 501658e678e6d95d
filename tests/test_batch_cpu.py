"""Batch descriptor validation, checked without a GPU: invalid problems are
rejected with a status before anything is launched."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def env():
    from exoblas_amd import lib, native
    return lib.load(), native, lib


def _op(native, **kw):
    ops = np.zeros(1, native.OP_DTYPE)
    base = dict(kind=native.OP_GEMM, trans_a=111, trans_b=111, m=4, n=4, k=4, a=16, lda=4, b=32, ldb=4, c=64,
                ldc=4, alpha=1.0, beta=0.0)
    base.update(kw)
    for k, v in base.items():
        ops[k] = v
    return ops


@pytest.mark.parametrize("kw,status", [
    (dict(trans_a=115), 1), (dict(m=-1), 2), (dict(lda=3), 4), (dict(ldc=2), 4), (dict(c=0), 8),
    (dict(kind=9), 1), (dict(kind=2, trans_a=112), 7), (dict(kind=2, ldb=0), 3),
    (dict(kind=3, trans_b=121, diag=131, ldc=2), 7), (dict(kind=3, trans_b=120, diag=131, ldc=1), 1),
])
def test_invalid_ops_rejected(env, kw, status):
    so, native, lib = env
    ops = _op(native, **kw)
    assert so.exoblas_batch_run(None, ops.ctypes.data, 1) == status
    lib.clear_error()


def test_empty_and_degenerate_batches(env):
    so, native, lib = env
    ops = _op(native, m=0)  # nothing to do: no launch
    assert so.exoblas_batch_run(None, ops.ctypes.data, 1) == 0
    assert so.exoblas_batch_run(None, None, 0) == 0


def test_tuning_table_roundtrip(env):
    so, native, lib = env
    t = (ctypes.c_int32 * 72)()
    assert so.exoblas_batch_get_tuning(t, 72) == 0
    saved = list(t)
    assert all(v in (0, 1, 2) for v in saved)
    bad = (ctypes.c_int32 * 72)(*([3] * 72))
    assert so.exoblas_batch_set_tuning(bad, 72) != 0
    assert so.exoblas_batch_set_tuning(t, 71) != 0
    lib.clear_error()
    new = (ctypes.c_int32 * 72)(*([1] * 72))
    assert so.exoblas_batch_set_tuning(new, 72) == 0
    out = (ctypes.c_int32 * 72)()
    so.exoblas_batch_get_tuning(out, 72)
    assert list(out) == [1] * 72
    assert so.exoblas_batch_set_tuning((ctypes.c_int32 * 72)(*saved), 72) == 0
