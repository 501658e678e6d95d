"""Serving loop failure paths and the pooled request layout.

* A queued request can be cancelled: it never runs, its output is untouched,
  its status reads cancelled; requests around it complete correctly.
* A malformed request is rejected on its own at submission with its status;
  the rest of the stream is unaffected.
* A batch whose bounded waits expire (forced from the host with a zero spin
  limit, so every kernel still ends normally) reports its requests failed,
  and the loop serves the next requests correctly once the limit is back.
* With every request reading its own window of large pools (the layout the
  serving benchmark uses), batched and per-call results match the oracle.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, NON_UNIT, Oracle
from util import assert_bound

pytestmark = pytest.mark.gpu

CANCELLED, HIP_ERROR, INVALID_LD, NOT_SUPPORTED = 10, 5, 4, 7


@pytest.fixture(scope="module")
def env():
    import torch
    from exoblas_amd import lib as L
    from exoblas_amd import native as N
    return torch, L.load(), L, N


def _op(N, **kw):
    op = N.BatchOp()
    for k, v in kw.items():
        setattr(op, k, v)
    return op


def _wait(so, srv, rid, limit_s=30.0):
    t0 = time.time()
    while True:
        st = so.exoblas_server_done(srv, rid)
        if st != 0:
            return st
        assert time.time() - t0 < limit_s, f"request {rid} did not finish"
        time.sleep(1e-4)


def _server(so, mode=0, streams=1, max_batch=1, budget=64 << 30):
    srv = ctypes.c_void_p()
    assert so.exoblas_server_create(ctypes.byref(srv), budget, streams, mode, max_batch) == 0
    return srv


def _gemv_op(N, A, x, y, m, n):
    return _op(N, kind=N.OP_GEMV, trans_a=NO_TRANS, m=m, n=n, a=A.data_ptr(), lda=n, b=x.data_ptr(), ldb=1,
               c=y.data_ptr(), ldc=1, alpha=1.5, beta=0.5)


@pytest.mark.parametrize("mode", [0, 1])
def test_cancel_queued_requests(env, mode):
    torch, so, L, N = env
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    big = 4096
    A = torch.empty((big, big), device="cuda").uniform_(-1, 1, generator=g)
    B = torch.empty((big, big), device="cuda").uniform_(-1, 1, generator=g)
    C = torch.zeros((big, big), device="cuda")
    m, n, count = 300, 500, 20
    Av = torch.empty((m, n), device="cuda").uniform_(-1, 1, generator=g)
    xv = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    Y = torch.empty((count, m), device="cuda").uniform_(-1, 1, generator=g)
    Y0 = Y.clone()
    srv = _server(so, mode=mode, streams=1, max_batch=1)
    try:
        # one long request holds the only slot while the small ones queue
        head = _op(N, kind=N.OP_GEMM, trans_a=NO_TRANS, trans_b=NO_TRANS, m=big, n=big, k=big, a=A.data_ptr(),
                   lda=big, b=B.data_ptr(), ldb=big, c=C.data_ptr(), ldc=big, alpha=1.0, beta=0.0)
        assert so.exoblas_server_submit(srv, ctypes.byref(head)) == 0
        ids = [so.exoblas_server_submit(srv, ctypes.byref(_gemv_op(N, Av, xv, Y[i], m, n))) for i in range(count)]
        assert ids == list(range(1, count + 1))
        cancel = ids[::2]
        got = [so.exoblas_server_cancel(srv, i) for i in cancel]
        assert got == [1] * len(cancel), got
        assert so.exoblas_server_cancel(srv, cancel[0]) == 0  # already cancelled
        assert so.exoblas_server_cancel(srv, 10 ** 6) < 0     # unknown id
        assert so.exoblas_server_drain(srv) == 0
        for i in ids:
            assert so.exoblas_server_done(srv, i) == (-CANCELLED if i in cancel else 1)
        assert so.exoblas_server_cancel(srv, ids[1]) == 0      # finished: too late
        st = N.ServerStats()
        so.exoblas_server_stats_get(srv, ctypes.byref(st))
        assert st.cancelled == len(cancel) and st.failed == 0
        assert st.completed == 1 + count - len(cancel)
    finally:
        so.exoblas_server_destroy(srv)
    torch.cuda.synchronize()
    Yh, Y0h = Y.cpu().numpy(), Y0.cpu().numpy()
    Ah, xh = Av.cpu().numpy(), xv.cpu().numpy()
    o = Oracle(np.float32)
    for j, i in enumerate(ids):
        if i in cancel:
            assert np.array_equal(Yh[j], Y0h[j]), f"cancelled request {i} wrote its output"
        else:
            want = Y0h[j].copy()
            o.gemv(NO_TRANS, m, n, 1.5, 0.5, Ah, n, xh, 1, want, 1)
            mag = 0.5 * np.abs(Y0h[j]) + 1.5 * (np.abs(Ah).astype(np.float64) @ np.abs(xh))
            assert_bound(Yh[j], want, mag, n, np.float32, what=f"request {i}")


def test_invalid_request_rejected_alone(env):
    torch, so, L, N = env
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    m, n = 200, 300
    Av = torch.empty((m, n), device="cuda").uniform_(-1, 1, generator=g)
    xv = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    Y = torch.empty((3, m), device="cuda").uniform_(-1, 1, generator=g)
    Y0 = Y.clone()
    srv = _server(so, mode=0, streams=2, max_batch=8)
    try:
        ok0 = so.exoblas_server_submit(srv, ctypes.byref(_gemv_op(N, Av, xv, Y[0], m, n)))
        bad = _gemv_op(N, Av, xv, Y[1], m, n)
        bad.lda = n - 1  # leading dimension below the row length
        L.clear_error()
        rb = so.exoblas_server_submit(srv, ctypes.byref(bad))
        assert rb == -INVALID_LD
        assert L.last_status() == INVALID_LD and "leading dimension" in L.last_error()
        L.clear_error()
        strided_trsv = _op(N, kind=N.OP_TRSV, trans_a=NO_TRANS, trans_b=LOWER, diag=NON_UNIT, m=64, a=Av.data_ptr(),
                           lda=n, c=Y[1].data_ptr(), ldc=2)
        assert so.exoblas_server_submit(srv, ctypes.byref(strided_trsv)) == -NOT_SUPPORTED
        L.clear_error()
        ok2 = so.exoblas_server_submit(srv, ctypes.byref(_gemv_op(N, Av, xv, Y[2], m, n)))
        assert ok0 == 0 and ok2 == 3
        assert so.exoblas_server_drain(srv) == 0
        assert so.exoblas_server_done(srv, 1) == -HIP_ERROR  # rejected
        st = N.ServerStats()
        so.exoblas_server_stats_get(srv, ctypes.byref(st))
        assert st.rejected == 2 and st.completed == 2 and st.failed == 0
    finally:
        so.exoblas_server_destroy(srv)
    torch.cuda.synchronize()
    assert torch.equal(Y[1], Y0[1])
    Ah, xh, Y0h, Yh = Av.cpu().numpy(), xv.cpu().numpy(), Y0.cpu().numpy(), Y.cpu().numpy()
    for j in (0, 2):
        want = Y0h[j].copy()
        Oracle(np.float32).gemv(NO_TRANS, m, n, 1.5, 0.5, Ah, n, xh, 1, want, 1)
        assert_bound(Yh[j], want, 0.5 * np.abs(Y0h[j]) + 1.5 * (np.abs(Ah) @ np.abs(xh)), n, np.float32,
                     what=f"valid request {j}")


@pytest.mark.parametrize("mode", [0, 1])
def test_expired_wait_fails_batch_then_recovers(env, mode):
    torch, so, L, N = env
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    n = 4096  # 64 dependent blocks: later ones always wait on earlier ones
    T = torch.empty((n, n), device="cuda").uniform_(-1, 1, generator=g) / n
    T.diagonal().copy_(n + torch.empty(n, device="cuda").uniform_(0, 1, generator=g))
    b = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    X = b.repeat(2, 1)

    def trsv(j):
        return _op(N, kind=N.OP_TRSV, trans_a=NO_TRANS, trans_b=LOWER, diag=NON_UNIT, m=n, a=T.data_ptr(), lda=n,
                   c=X[j].data_ptr(), ldc=1)
    srv = _server(so, mode=mode, streams=1, max_batch=4)
    try:
        assert so.exoblas_set_spin_limit(0) == 0
        try:
            r0 = so.exoblas_server_submit(srv, ctypes.byref(trsv(0)))
            assert _wait(so, srv, r0) == -HIP_ERROR
        finally:
            assert so.exoblas_set_spin_limit(-1) == 0
        r1 = so.exoblas_server_submit(srv, ctypes.byref(trsv(1)))
        assert _wait(so, srv, r1) == 1
        st = N.ServerStats()
        so.exoblas_server_stats_get(srv, ctypes.byref(st))
        assert st.failed == 1 and st.completed == 2
    finally:
        so.exoblas_server_destroy(srv)
    L.clear_error()
    torch.cuda.synchronize()
    want = b.cpu().numpy().copy()
    Oracle(np.float32).trsv(LOWER, NO_TRANS, NON_UNIT, n, want, 1, T.cpu().numpy(), n)
    assert_bound(X[1].cpu().numpy(), want, np.abs(want) + 1, n, np.float32, what="trsv after recovery")


def test_pooled_layout_matches_oracle(env):
    torch, so, L, N = env
    from exoblas_amd import serving
    s = serving.make_stream(300, rate=50000.0, seed=6)
    s.m[:] = np.minimum(s.m, 700)
    s.n[:] = np.minimum(s.n, 700)
    s.k[:] = np.minimum(s.k, 700)
    s.n[s.kind == N.OP_TRSV] = s.m[s.kind == N.OP_TRSV]
    w = serving.Workload(s, pool_bytes=256 << 20, tri_pool=4)
    assert len({int(a) for a in w.ops["a"][s.kind == N.OP_GEMM]}) > 1  # distinct windows
    results = {}
    for mode in (0, 1):
        w.reset()
        torch.cuda.synchronize()
        stats, _ = serving.run(w, mode=mode, streams=3, max_batch=64)
        assert stats["completed"] == len(s.kind) and stats["failed"] == 0
        results[mode] = [w.output(i).cpu().numpy().copy() for i in range(len(s.kind))]
    o = Oracle(np.float32)
    init = w.init.cpu().numpy()
    for i in range(0, len(s.kind), 5):
        m, n, k = int(s.m[i]), int(s.n[i]), int(s.k[i])
        off = int(w.out_off[i])
        ops = w.operands(i)
        if s.kind[i] == N.OP_GEMM:
            Am, Bm = ops
            c0 = init[off:off + m * n].reshape(m, n)
            want = c0.copy()
            o.mscal(m, n, 0.5, want, n)
            o.gemm(NO_TRANS, NO_TRANS, m, n, k, 1.0, Am, k, Bm, n, want, n)
            mag, K = 0.5 * np.abs(c0) + np.abs(Am).astype(np.float64) @ np.abs(Bm), k
        elif s.kind[i] == N.OP_GEMV:
            Am, xv = ops
            y0 = init[off:off + m]
            want = y0.copy()
            o.gemv(NO_TRANS, m, n, 1.0, 0.5, Am, n, xv, 1, want, 1)
            mag, K = 0.5 * np.abs(y0) + np.abs(Am).astype(np.float64) @ np.abs(xv), n
        else:
            (Tm,) = ops
            want = init[off:off + m].copy()
            o.trsv(LOWER, NO_TRANS, NON_UNIT, m, want, 1, Tm, m)
            mag, K = np.abs(want) + 1, m
        for mode in (0, 1):
            assert_bound(results[mode][i].reshape(want.shape), want, mag, K, np.float32,
                         what=f"pooled request {i} mode {mode}")
