"""exoblas_sdot_noalloc_async against exoblas_sdot_stride_any_async, the entry
point it is a variant of: same kernel and launch shape, so the results must be
bit-identical, on two calls in a row (the second smaller and strided, so the
fixed scratch keeps the first call's contents), plus a float64 bound, the
argument checks, and hipGraph capture and replay (the callers it is for)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def env():
    import torch
    from exoblas_amd import lib as L
    so = L.load()
    return torch, so, L


def _win(L, t, n, inc):
    """Window at element 0 of an n-vector at stride inc over tensor t (a
    negative stride walks down from the last position of t's range)."""
    base = t.data_ptr() + ((n - 1) * -inc * 4 if inc < 0 and n > 0 else 0)
    return L.win1(base, inc)


def _run(so, stream, fn, n, wx, wy, res):
    return fn(stream, n, wx, wy, ctypes.c_void_p(res.data_ptr()))


@pytest.mark.parametrize("first,second", [((1 << 22) + 37, (1000, -2, 3)), ((1 << 24), (4099, 1, 1))])
def test_noalloc_matches_sdot(env, first, second):
    torch, so, L = env
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    stream = torch.cuda.current_stream().cuda_stream
    calls = [(first, 1, 1), second]
    for n, incx, incy in calls:
        x = torch.empty(1 + (n - 1) * abs(incx), device="cuda").uniform_(-1, 1, generator=g)
        y = torch.empty(1 + (n - 1) * abs(incy), device="cuda").uniform_(-1, 1, generator=g)
        r_ref = torch.full((1,), float("nan"), device="cuda")
        r_na = torch.full((1,), float("nan"), device="cuda")
        wx, wy = _win(L, x, n, incx), _win(L, y, n, incy)
        assert _run(so, stream, so.exoblas_sdot_stride_any_async, n, wx, wy, r_ref) == 0
        assert _run(so, stream, so.exoblas_sdot_noalloc_async, n, wx, wy, r_na) == 0
        torch.cuda.synchronize()
        assert r_na.cpu().numpy().tobytes() == r_ref.cpu().numpy().tobytes(), (n, r_na.item(), r_ref.item())
        xs = x[::abs(incx)].double()
        ys = y[::abs(incy)].double()
        if incx < 0:
            xs = xs.flip(0)
        if incy < 0:
            ys = ys.flip(0)
        want = float(xs @ ys)
        mag = float((xs * ys).abs().sum())
        assert abs(r_na.item() - want) <= 4 * (n + 2) * EPS32 * mag, (n, r_na.item(), want)


def test_noalloc_rejects_bad_arguments(env):
    torch, so, L = env
    x = torch.ones(16, device="cuda")
    res = torch.zeros(1, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    w = L.win1(x.data_ptr(), 1)
    assert so.exoblas_sdot_noalloc_async(stream, -1, w, w, ctypes.c_void_p(res.data_ptr())) == 2
    assert so.exoblas_sdot_noalloc_async(stream, 16, w, L.win1(0, 1), ctypes.c_void_p(res.data_ptr())) == 8
    assert so.exoblas_sdot_noalloc_async(stream, 16, w, w, ctypes.c_void_p(0)) == 8
    so.exoblas_clear_error()
    res.fill_(7.0)
    assert so.exoblas_sdot_noalloc_async(stream, 0, w, w, ctypes.c_void_p(res.data_ptr())) == 0
    torch.cuda.synchronize()
    assert res.item() == 0.0


def test_noalloc_graph_replay(env):
    """Captured once, replayed with new inputs written eagerly between
    replays: every replay equals an eager call on those inputs."""
    torch, so, L = env
    g = torch.Generator(device="cuda")
    n = (1 << 20) + 3
    x = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    y = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    res = torch.zeros(1, device="cuda")
    ref = torch.zeros(1, device="cuda")
    wx, wy = L.win1(x.data_ptr(), 1), L.win1(y.data_ptr(), 1)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        assert so.exoblas_sdot_noalloc_async(torch.cuda.current_stream().cuda_stream, n, wx, wy,
                                             ctypes.c_void_p(res.data_ptr())) == 0
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(4):
        x.uniform_(-1, 1, generator=g)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert so.exoblas_sdot_stride_1_async(stream, n, wx, wy, ctypes.c_void_p(ref.data_ptr())) == 0
        torch.cuda.synchronize()
        assert res.cpu().numpy().tobytes() == ref.cpu().numpy().tobytes(), (rep, res.item(), ref.item())
