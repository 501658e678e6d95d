"""Runtime behaviour on the GPU: bounded waits that expire are reported, grown
buffers stay valid for graphs captured before the growth, and growth during a
capture is refused instead of allocating.

The timeout tests set the wait bound to zero polls from the host (a test hook),
so waits give up at their first miss: every kernel still ends normally, the
results are garbage, and the status must say so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, NON_UNIT, Oracle
from util import assert_bound

pytestmark = pytest.mark.gpu

TIMEOUT = 9
NOT_SUPPORTED = 7


@pytest.fixture(scope="module")
def env():
    import torch
    from exoblas_amd import lib as L
    so = L.load()
    return torch, so, L


def dd_matrix(rng, n):
    A = (rng.uniform(-1, 1, (n, n)) / n).astype(np.float32)
    A[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(np.float32)
    return A


def test_timeout_reported_async(env):
    torch, so, L = env
    rng = np.random.default_rng(0)
    n = 4096  # 64 dependent blocks: later blocks always wait on earlier ones
    A = dd_matrix(rng, n)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    dA, dx = torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda()
    s = torch.cuda.Stream()

    def solve():
        dx.copy_(torch.from_numpy(b))
        torch.cuda.synchronize()
        st = so.exoblas_strsv_rm_stride_1_async(s.cuda_stream, LOWER, NO_TRANS, NON_UNIT, n,
                                                L.win1(dx.data_ptr(), 1), L.win2(dA.data_ptr(), n))
        assert st == 0
        return so.exoblas_stream_check(s.cuda_stream)

    assert solve() == 0
    try:
        assert so.exoblas_set_spin_limit(0) == 0
        assert solve() == TIMEOUT
        assert so.exoblas_stream_check(s.cuda_stream) == 0  # the check clears the word
    finally:
        assert so.exoblas_set_spin_limit(-1) == 0
    assert solve() == 0
    xo = b.copy()
    Oracle(np.float32).trsv(LOWER, NO_TRANS, NON_UNIT, n, xo, 1, A, n)
    assert_bound(dx.cpu().numpy(), xo, np.abs(xo) + 1, n, np.float32, what="trsv after the timeout test")


def test_timeout_reported_sync(env):
    """The synchronous drop-in path checks the word itself and reports
    EXOBLAS_STATUS_TIMEOUT through the thread's status."""
    torch, so, L = env
    rng = np.random.default_rng(1)
    n = 4096
    A = dd_matrix(rng, n)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    try:
        assert so.exoblas_set_spin_limit(0) == 0
        so.exoblas_clear_error()
        so.exo_strsv_rm_stride_1(None, LOWER, NO_TRANS, NON_UNIT, n, L.win1(x.ctypes.data, 1),
                                 L.win2(A.ctypes.data, n))
        assert so.exoblas_last_status() == TIMEOUT
    finally:
        so.exoblas_set_spin_limit(-1)
        so.exoblas_clear_error()
    x = rng.uniform(-1, 1, n).astype(np.float32)
    xo = x.copy()
    so.exo_strsv_rm_stride_1(None, LOWER, NO_TRANS, NON_UNIT, n, L.win1(x.ctypes.data, 1), L.win2(A.ctypes.data, n))
    L.check()
    Oracle(np.float32).trsv(LOWER, NO_TRANS, NON_UNIT, n, xo, 1, A, n)
    assert_bound(x, xo, np.abs(xo) + 1, n, np.float32, what="trsv sync after the timeout test")


def test_graph_replay_after_growth(env):
    """A dot captured at a small size, then an eager call large enough to grow
    the stream's flag area, then replays: the graph still reads and writes the
    area it was captured with (retired, not freed)."""
    torch, so, L = env
    s = torch.cuda.Stream()
    g = torch.Generator(device="cuda")
    small, big = 1 << 16, 1 << 26
    x = torch.empty(big, device="cuda").uniform_(-1, 1, generator=g)
    y = torch.empty(big, device="cuda").uniform_(-1, 1, generator=g)
    res = torch.zeros(1, device="cuda")
    res_big = torch.zeros(1, device="cuda")

    def dot(stream, n, out):
        return so.exoblas_sdot_stride_1_async(stream, n, L.win1(x.data_ptr(), 1), L.win1(y.data_ptr(), 1),
                                              ctypes.c_void_p(out.data_ptr()))
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        assert dot(s.cuda_stream, small, res) == 0
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        assert dot(torch.cuda.current_stream().cuda_stream, small, res) == 0
    # grows the flag area (2 workgroups per CU publish their partials)
    with torch.cuda.stream(s):
        assert dot(s.cuda_stream, big, res_big) == 0
    s.synchronize()
    eps = float(np.finfo(np.float32).eps)
    for rep in range(3):
        x[:small].uniform_(-1, 1, generator=g)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        want = float(x[:small].double() @ y[:small].double())
        mag = float((x[:small].double() * y[:small].double()).abs().sum())
        assert abs(res.item() - want) <= 4 * small * eps * mag, rep
        with torch.cuda.stream(s):  # eager work between replays
            assert dot(s.cuda_stream, big, res_big) == 0
        s.synchronize()
    want = float(x.double() @ y.double())
    assert abs(res_big.item() - want) <= 4 * big * eps * float((x.double() * y.double()).abs().sum())


def _hip():
    """The HIP runtime this process already uses (torch's copy), by the path it
    is mapped from, so no second runtime is loaded."""
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert len(paths) == 1, paths
    hip = ctypes.CDLL(paths.pop())
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_growth_during_capture_refused(env):
    """On a stream the library has never seen: its first call during a
    capture is refused (its state would need allocating); after one small
    eager call, a call that would grow the stream's flag area during a capture
    is refused too; both return EXOBLAS_STATUS_NOT_SUPPORTED and enqueue
    nothing, and the same calls then work eagerly."""
    torch, so, L = env
    hip = _hip()
    raw = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(raw)) == 0
    # HIP may recycle the handle of a stream destroyed earlier in this
    # process; releasing it drops any state the library kept for that one
    assert so.exoblas_stream_release(raw.value) == 0
    try:
        s = torch.cuda.ExternalStream(raw.value)
        big = 1 << 26
        x = torch.ones(big, device="cuda")
        res = torch.zeros(1, device="cuda")

        def dot(stream, n):
            return so.exoblas_sdot_stride_1_async(stream, n, L.win1(x.data_ptr(), 1), L.win1(x.data_ptr(), 1),
                                                  ctypes.c_void_p(res.data_ptr()))
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=s):
            st_first = dot(raw.value, 1 << 12)
        assert st_first == NOT_SUPPORTED
        so.exoblas_clear_error()
        assert dot(raw.value, 1 << 12) == 0  # the stream's state, smallest flag area
        assert so.exoblas_stream_check(raw.value) == 0
        assert res.item() == float(1 << 12)
        graph2 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph2, stream=s):
            st_grow = dot(raw.value, big)
        assert st_grow == NOT_SUPPORTED
        so.exoblas_clear_error()
        assert dot(raw.value, big) == 0
        assert so.exoblas_stream_check(raw.value) == 0
        assert res.item() == float(big)
        del graph, graph2
    finally:
        torch.cuda.synchronize()
        assert so.exoblas_stream_release(raw.value) == 0
        hip.hipStreamDestroy(raw)


def test_stream_release(env):
    """exoblas_stream_release frees a stream's state: the next call on the
    same handle builds it afresh (and gives the same result), releasing an
    unknown or already released stream is a no-op, and a released stream can
    be destroyed while later calls on other streams still work."""
    torch, so, L = env
    hip = _hip()
    raw = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(raw)) == 0
    try:
        n = 1 << 22
        x = torch.ones(n, device="cuda")
        res = torch.zeros(1, device="cuda")
        torch.cuda.synchronize()
        for _ in range(2):
            assert so.exoblas_sdot_stride_1_async(raw.value, n, L.win1(x.data_ptr(), 1), L.win1(x.data_ptr(), 1),
                                                  ctypes.c_void_p(res.data_ptr())) == 0
            assert so.exoblas_stream_check(raw.value) == 0
            assert res.item() == float(n)
            res.zero_()
            torch.cuda.synchronize()
            assert so.exoblas_stream_release(raw.value) == 0
            assert so.exoblas_stream_release(raw.value) == 0  # already released
    finally:
        torch.cuda.synchronize()
        hip.hipStreamDestroy(raw)
    s = torch.cuda.Stream()
    assert so.exoblas_sdot_stride_1_async(s.cuda_stream, 1000, L.win1(x.data_ptr(), 1), L.win1(x.data_ptr(), 1),
                                          ctypes.c_void_p(res.data_ptr())) == 0
    s.synchronize()
    assert res.item() == 1000.0
