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


def test_queue_order(env):
    """The work queue of a mixed batch: every item of every problem exactly
    once, each trsv chain's steps in order, one step per chain per round
    with gemm / gemv items between the rounds (host-side only)."""
    so, native, lib = env
    rng = np.random.default_rng(0)
    specs = []
    for i in range(60):
        r = i % 4
        if r == 0:
            n = int(rng.integers(1, 2500))
            specs.append(dict(kind=native.OP_TRSV, trans_a=111, trans_b=122, diag=131, m=n, n=n, k=0, lda=n,
                              ldc=1, b=0))
        elif r == 3:
            m, n = (int(v) for v in rng.integers(1, 3000, 2))
            specs.append(dict(kind=native.OP_GEMV, m=m, n=n, k=0, lda=n, ldb=1, ldc=1))
        else:
            m, n, k = (int(v) for v in rng.integers(1, 1100, 3))
            specs.append(dict(kind=native.OP_GEMM, m=m, n=n, k=k, lda=k, ldb=n, ldc=n))
    ops = np.concatenate([_op(native, **s) for s in specs])
    n_items, n_prob = ctypes.c_int64(), ctypes.c_int64()
    assert so.exoblas_batch_queue(ops.ctypes.data, len(ops), None, None, 0, ctypes.byref(n_items),
                                  ctypes.byref(n_prob)) == 0
    cap = max(n_items.value, n_prob.value)
    order = np.zeros(cap, np.uint64)
    kinds = np.zeros(cap, np.int32)
    assert so.exoblas_batch_queue(ops.ctypes.data, len(ops), order.ctypes.data, kinds.ctypes.data, cap,
                                  ctypes.byref(n_items), ctypes.byref(n_prob)) == 0
    order, kinds = order[:n_items.value], kinds[:n_prob.value]
    assert n_prob.value == len(ops)
    prob, local = (order >> np.uint64(32)).astype(np.int64), (order & np.uint64(0xffffffff)).astype(np.int64)
    assert prob.max() < n_prob.value
    trsv_sizes = sorted((s["m"] for s in specs if s["kind"] == native.OP_TRSV), reverse=True)
    chains = [p for p in range(n_prob.value) if kinds[p] == native.OP_TRSV]
    assert sorted(-(-n // 64) for n in trsv_sizes) == sorted(int((prob == p).sum()) for p in chains)
    for p in range(n_prob.value):
        loc = local[prob == p]
        assert sorted(loc.tolist()) == list(range(len(loc))), p  # every item exactly once
        if kinds[p] == native.OP_TRSV:
            assert (np.diff(loc) > 0).all()  # steps in chain order
    # round r holds step r of every chain longer than r, then other items
    is_trsv = np.isin(prob, chains)
    pos = np.flatnonzero(is_trsv)
    rounds = local[pos]
    assert (np.diff(rounds) >= 0).all()
    gaps = np.flatnonzero(np.diff(rounds) > 0)
    assert all(not is_trsv[pos[g] + 1: pos[g + 1]].all() for g in gaps) and len(gaps) > 0
