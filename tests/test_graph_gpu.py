"""Single-launch reductions and split gemv-T captured in a hipGraph: their
publication markers come from a per-stream counter on the device, so every
replay must reproduce the eager result for the inputs of that replay (the
inputs are rewritten eagerly between replays)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    from exoblas_amd import lib as L
    so = L.load()
    return torch, so, L


def test_dot_and_gemv_t_graph_replay(env):
    torch, so, L = env
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    n = (1 << 20) + 37
    M = N = 4096  # gemv T over 16 column blocks: row splits, last-workgroup sums
    x = torch.empty(n, device="cuda")
    y = torch.empty(n, device="cuda")
    A = torch.empty((M, N), device="cuda")
    gx = torch.empty(M, device="cuda")
    gy = torch.empty(N, device="cuda")
    res = torch.zeros(1, device="cuda")
    ga, gb = ctypes.byref(ctypes.c_float(0.5)), ctypes.byref(ctypes.c_float(1.5))
    s = torch.cuda.Stream()

    def fill(seed):
        g.manual_seed(seed)
        for t in (x, y, A, gx, gy):
            t.uniform_(-1, 1, generator=g)

    def launch(stream):
        assert so.exoblas_sdot_stride_1_async(stream, n, L.win1(x.data_ptr(), 1), L.win1(y.data_ptr(), 1),
                                              ctypes.c_void_p(res.data_ptr())) == 0
        assert so.exoblas_sgemv_rm_stride_1_async(stream, 112, N, M, ga, gb, L.win2(A.data_ptr(), N),
                                                  L.win2(A.data_ptr(), N), L.win1(gx.data_ptr(), 1),
                                                  L.win1(gy.data_ptr(), 1)) == 0

    def expect():
        x64, y64 = x.double().cpu().numpy(), y.double().cpu().numpy()
        dot = float(x64 @ y64)
        gy_want = 1.5 * gy.double().cpu().numpy() + 0.5 * (A.double().cpu().numpy().T @ gx.double().cpu().numpy())
        return dot, gy_want

    # eager warm-up (grows the stream's buffers outside the capture)
    fill(0)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        launch(s.cuda_stream)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    fill(1)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        launch(torch.cuda.current_stream().cuda_stream)
    for rep in range(4):
        fill(100 + rep)
        gy0 = gy.clone()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got_gy = gy.double().cpu().numpy()
        gy.copy_(gy0)
        dot, gy_want = expect()
        eps = float(np.finfo(np.float32).eps)
        mag = float(np.abs(x.double().cpu().numpy() * y.double().cpu().numpy()).sum())
        assert abs(res.item() - dot) <= 4 * n * eps * mag, f"replay {rep}: dot"
        gmag = 1.5 * np.abs(gy0.double().cpu().numpy()) + 0.5 * (np.abs(A.double().cpu().numpy()).T @
                                                              np.abs(gx.double().cpu().numpy()))
        assert np.all(np.abs(got_gy - gy_want) <= 2 * M * eps * gmag + 1e-30), f"replay {rep}: gemv T"
        # and the replay equals an eager launch on the same inputs
        res_g = res.clone()
        with torch.cuda.stream(s):
            launch(s.cuda_stream)
        s.synchronize()
        assert torch.equal(res, res_g), f"replay {rep}: dot differs from the eager launch"
        assert np.array_equal(gy.double().cpu().numpy(), got_gy), f"replay {rep}: gemv T differs from eager"
        gy.copy_(gy0)


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_trsv_graph_replay(env, dt):
    """A captured triangular solve (look-back chain with tagged x blocks)
    replayed with a new right-hand side each time equals the eager solve of
    that right-hand side, bit for bit, and solves the system."""
    torch, so, L = env
    tdt = getattr(torch, dt)
    p = "s" if dt == "float32" else "d"
    n = 1000
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    A = torch.empty((n, n), dtype=tdt, device="cuda").uniform_(-1, 1, generator=g) / n
    A.diagonal().copy_(n + torch.empty(n, dtype=tdt, device="cuda").uniform_(0, 1, generator=g))
    x = torch.empty(n, dtype=tdt, device="cuda")
    f = getattr(so, f"exoblas_{p}trsv_rm_stride_1_async")
    s = torch.cuda.Stream()

    def launch(stream):
        assert f(stream, 122, 111, 131, n, L.win1(x.data_ptr(), 1), L.win2(A.data_ptr(), n)) == 0

    x.uniform_(-1, 1, generator=g)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        launch(s.cuda_stream)  # eager warm-up outside the capture
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        launch(torch.cuda.current_stream().cuda_stream)
    L64 = torch.tril(A).double().cpu().numpy()
    for rep in range(4):
        b = torch.empty(n, dtype=tdt, device="cuda").uniform_(-1, 1, generator=g)
        x.copy_(b)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = x.clone()
        x.copy_(b)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            launch(s.cuda_stream)
        s.synchronize()
        assert torch.equal(got, x), f"replay {rep} differs from the eager solve"
        r = L64 @ got.double().cpu().numpy() - b.double().cpu().numpy()
        eps = float(np.finfo(np.dtype(dt)).eps)
        assert np.abs(r).max() <= 8 * n * eps, f"replay {rep}: residual {np.abs(r).max()}"


def test_banded_syr2k_trmm_graph_replay(env):
    """sbmv, tbmv, tbsv (both solvers), syr2k and strmm captured in one hipGraph
    (their scratch is grown by an eager call first): each replay on freshly
    written inputs equals the eager calls on the same inputs, bit for bit."""
    torch, so, L = env
    g = torch.Generator(device="cuda")
    n, k, kw = 20000, 16, 40
    N2, K2 = 512, 256
    M3, N3 = 1024, 777
    f = lambda: ctypes.byref(ctypes.c_float(1.25))
    A = torch.empty((n, k + 1), device="cuda")
    Aw = torch.empty((n, kw + 1), device="cuda")
    x, y, xt, xs, xw = (torch.empty(n, device="cuda") for _ in range(5))
    Bk, Ck = torch.empty((N2, K2), device="cuda"), torch.empty((N2, N2), device="cuda")
    Ak = torch.empty((N2, K2), device="cuda")
    At, Bt = torch.empty((M3, M3), device="cuda"), torch.empty((M3, N3), device="cuda")
    outs = (y, xt, xs, xw, Ck, Bt)
    s = torch.cuda.Stream()

    def fill(seed):
        g.manual_seed(seed)
        for t in (A, Aw, x, y, xt, xs, xw, Bk, Ck, Ak, At, Bt):
            t.uniform_(-1, 1, generator=g)
        A[:, 0] += k + 1  # dominant diagonals for the solves (upper storage: column 0)
        Aw[:, 0] += kw + 1

    def launch(st):
        assert so.exoblas_ssbmv_row_major_Upper_stride_1_async(st, n, k, f(), L.win2(A.data_ptr(), k + 1),
                                                               L.win1(x.data_ptr(), 1), L.win1(y.data_ptr(), 1)) == 0
        assert so.exoblas_stbmv_row_major_Upper_Trans_stride_1_async(st, n, k, L.win1(xt.data_ptr(), 1),
                                                                     L.win2(A.data_ptr(), k + 1), 0) == 0
        assert so.exoblas_stbsv_row_major_Upper_NonTrans_stride_1_async(st, n, k, L.win1(xs.data_ptr(), 1),
                                                                        L.win2(A.data_ptr(), k + 1), 0) == 0
        assert so.exoblas_stbsv_row_major_Upper_Trans_stride_any_async(st, n, kw, L.win1(xw.data_ptr(), 1),
                                                                       L.win2(Aw.data_ptr(), kw + 1), 0) == 0
        wa, wb = L.win2(Ak.data_ptr(), K2), L.win2(Bk.data_ptr(), K2)
        assert so.exoblas_ssyr2k_rm_stride_1_async(st, 122, 111, N2, K2, f(), wa, wa, wb, wb, wa, wa, wb, wb,
                                                   L.win2(Ck.data_ptr(), N2)) == 0
        assert so.exoblas_strmm_row_major_Left_Lower_NonTrans_async(st, M3, N3, f(), L.win2(At.data_ptr(), M3),
                                                                    L.win2(Bt.data_ptr(), N3), 0) == 0

    fill(0)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        launch(s.cuda_stream)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    fill(1)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        launch(torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        fill(200 + rep)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [t.clone() for t in outs]
        fill(200 + rep)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            launch(s.cuda_stream)
        s.synchronize()
        for name, a, b in zip(("sbmv", "tbmv", "tbsv", "tbsv wide", "syr2k", "trmm"), got, outs):
            assert torch.equal(a, b), f"replay {rep}: {name} differs from the eager call"
        assert bool(torch.isfinite(xs).all()) and bool(torch.isfinite(xw).all())


@pytest.mark.parametrize("mnk", [(512, 512, 512), (64, 96, 257)])
def test_small_gemm_one_launch_split_graph_replay(env, mnk):
    """Small outputs with a long K split K over 64 x 64 tiles (slices in the
    stream's scratch, summed in slice order by a second launch): a captured
    call replayed on fresh inputs equals the eager call bit for bit, and the
    product matches float64."""
    torch, so, L = env
    M, N, K = mnk
    g = torch.Generator(device="cuda")
    A, B, Cm = (torch.empty(sh, device="cuda") for sh in ((M, K), (K, N), (M, N)))
    al = ctypes.byref(ctypes.c_float(1.5))
    s = torch.cuda.Stream()

    def fill(seed):
        g.manual_seed(seed)
        for t in (A, B, Cm):
            t.uniform_(-1, 1, generator=g)

    def launch(st):
        assert so.exoblas_sgemm_stride_1_async(st, 111, 111, M, N, K, al, L.win2(A.data_ptr(), K),
                                               L.win2(B.data_ptr(), N), L.win2(A.data_ptr(), K),
                                               L.win2(B.data_ptr(), N), L.win2(Cm.data_ptr(), N)) == 0

    fill(0)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        launch(s.cuda_stream)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    fill(1)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        launch(torch.cuda.current_stream().cuda_stream)
    eps = float(np.finfo(np.float32).eps)
    for rep in range(3):
        fill(300 + rep)
        C0 = Cm.double()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = Cm.clone()
        want = C0 + 1.5 * (A.double() @ B.double())
        mag = C0.abs() + 1.5 * (A.double().abs() @ B.double().abs())
        assert bool(((got.double() - want).abs() <= 4 * (K + 2) * eps * mag).all()), f"replay {rep}: product"
        fill(300 + rep)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            launch(s.cuda_stream)
        s.synchronize()
        assert torch.equal(got, Cm), f"replay {rep}: differs from the eager call"
