"""Batched persistent kernel, replayable plans and the serving loop.

Parity: every problem of a mixed batch (gemm in all trans combinations with
and without beta, gemv, trsv in all eight uplo/trans/diag variants, uneven
sizes) against the oracle.  Plans are captured in a hipGraph and replayed
several times with eager work between replays; every replay must equal the
eager result.  The serving loop must give each request the same result alone
and inside a mixed stream, keep in-flight operand bytes within its budget and
reject a request that can never fit.
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np
import pytest

from oracle import LOWER, NO_TRANS, NON_UNIT, TRANS, UNIT, UPPER, Oracle
from util import assert_bound

pytestmark = pytest.mark.gpu

EPS = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def env():
    import torch
    from exoblas_amd import lib, native
    so = lib.load()
    return torch, so, native


def _mixed_problems(rng):
    """(op kind, params) for a mixed batch, host arrays included."""
    probs = []
    for (M, N, K), ta, tb, beta in zip(
            [(16, 16, 16), (33, 70, 129), (200, 64, 1), (1, 1000, 17), (130, 129, 300), (512, 512, 64),
             (7, 3, 1024), (64, 64, 0)],
            itertools.cycle([NO_TRANS, TRANS]), itertools.cycle([NO_TRANS, NO_TRANS, TRANS, TRANS]),
            itertools.cycle([0.0, 0.5, 1.25])):
        A = rng.uniform(-1, 1, (M, K) if ta == NO_TRANS else (K, M)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N) if tb == NO_TRANS else (N, K)).astype(np.float32)
        C = rng.uniform(-1, 1, (M, N)).astype(np.float32)
        probs.append(("gemm", dict(M=M, N=N, K=K, ta=ta, tb=tb, A=A, B=B, C=C, alpha=1.5, beta=beta)))
    for m, n in [(64, 64), (100, 3000), (1, 5), (777, 129), (4096, 33)]:
        probs.append(("gemv", dict(m=m, n=n, A=rng.uniform(-1, 1, (m, n)).astype(np.float32),
                                   x=rng.uniform(-1, 1, n).astype(np.float32),
                                   y=rng.uniform(-1, 1, m).astype(np.float32), alpha=0.75, beta=0.5)))
    for (uplo, tr, dg), n in zip(itertools.product([UPPER, LOWER], [NO_TRANS, TRANS], [UNIT, NON_UNIT]),
                                 [33, 64, 65, 200, 1, 300, 511, 130]):
        T = (rng.uniform(-1, 1, (n, n)) / n).astype(np.float32)
        T[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(np.float32)
        probs.append(("trsv", dict(n=n, uplo=uplo, tr=tr, dg=dg, A=T, x=rng.uniform(-1, 1, n).astype(np.float32))))
    return probs


def _to_device(torch, native, probs):
    ops = np.zeros(len(probs), native.OP_DTYPE)
    keep = []
    for i, (kind, p) in enumerate(probs):
        o = ops[i]
        if kind == "gemm":
            dA, dB, dC = (torch.from_numpy(p[k]).cuda() for k in "ABC")
            keep += [dA, dB, dC]
            o["kind"], o["trans_a"], o["trans_b"] = native.OP_GEMM, p["ta"], p["tb"]
            o["m"], o["n"], o["k"] = p["M"], p["N"], p["K"]
            o["a"], o["lda"] = dA.data_ptr(), max(1, p["A"].shape[1])
            o["b"], o["ldb"] = dB.data_ptr(), max(1, p["B"].shape[1])
            o["c"], o["ldc"] = dC.data_ptr(), p["N"]
            o["alpha"], o["beta"] = p["alpha"], p["beta"]
            p["out"] = dC
        elif kind == "gemv":
            dA, dx, dy = torch.from_numpy(p["A"]).cuda(), torch.from_numpy(p["x"]).cuda(), torch.from_numpy(p["y"]).cuda()
            keep += [dA, dx, dy]
            o["kind"], o["trans_a"] = native.OP_GEMV, NO_TRANS
            o["m"], o["n"] = p["m"], p["n"]
            o["a"], o["lda"], o["b"], o["ldb"], o["c"], o["ldc"] = dA.data_ptr(), p["n"], dx.data_ptr(), 1, dy.data_ptr(), 1
            o["alpha"], o["beta"] = p["alpha"], p["beta"]
            p["out"] = dy
        else:
            dA, dx = torch.from_numpy(p["A"]).cuda(), torch.from_numpy(p["x"]).cuda()
            keep += [dA, dx]
            o["kind"], o["trans_a"], o["trans_b"], o["diag"] = native.OP_TRSV, p["tr"], p["uplo"], p["dg"]
            o["m"] = o["n"] = p["n"]
            o["a"], o["lda"], o["c"], o["ldc"] = dA.data_ptr(), p["n"], dx.data_ptr(), 1
            p["out"] = dx
    return ops, keep


def _check(probs):
    o = Oracle(np.float32)
    for kind, p in probs:
        got = p["out"].cpu().numpy()
        if kind == "gemm":
            want = p["C"].copy()
            o.mscal(p["M"], p["N"], p["beta"], want, p["N"])
            o.gemm(p["ta"], p["tb"], p["M"], p["N"], p["K"], p["alpha"], p["A"], max(1, p["A"].shape[1]), p["B"],
                   max(1, p["B"].shape[1]), want, p["N"])
            opA = np.abs(p["A"] if p["ta"] == NO_TRANS else p["A"].T).astype(np.float64)
            opB = np.abs(p["B"] if p["tb"] == NO_TRANS else p["B"].T).astype(np.float64)
            mag = abs(p["beta"]) * np.abs(p["C"]) + p["alpha"] * (opA @ opB)
            assert_bound(got, want, mag, p["K"], np.float32, what=f"batched gemm {p['M']}x{p['N']}x{p['K']}")
        elif kind == "gemv":
            want = p["y"].copy()
            o.gemv(NO_TRANS, p["m"], p["n"], p["alpha"], p["beta"], p["A"], p["n"], p["x"], 1, want, 1)
            mag = 0.5 * np.abs(p["y"]) + 0.75 * (np.abs(p["A"]).astype(np.float64) @ np.abs(p["x"]))
            assert_bound(got, want, mag, p["n"], np.float32, what=f"batched gemv {p['m']}x{p['n']}")
        else:
            want = p["x"].copy()
            o.trsv(p["uplo"], p["tr"], p["dg"], p["n"], want, 1, p["A"], p["n"])
            assert_bound(got, want, np.abs(want) + 1, p["n"], np.float32, what=f"batched trsv n={p['n']}")


def test_batch_run_mixed(env, rng):
    torch, so, native = env
    probs = _mixed_problems(rng)
    ops, keep = _to_device(torch, native, probs)
    st = so.exoblas_batch_run(torch.cuda.current_stream().cuda_stream, ops.ctypes.data, len(ops))
    assert st == 0
    torch.cuda.synchronize()
    _check(probs)


def test_batch_beta_zero_reads_output(env, rng):
    """beta = 0 with NaN / Inf already in C and y: the batch forms beta * C as
    the per-call path does (the reference's mscal, then the product), so those
    entries stay NaN in both; finite entries get alpha op(A) op(B) (gemm) and
    alpha A x (gemv)."""
    torch, so, native = env
    probs = []
    for M, N, K in [(70, 90, 33), (256, 256, 64)]:
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        C = rng.uniform(-1, 1, (M, N)).astype(np.float32)
        C[::7, ::5] = np.nan
        C[3::11, 2::9] = np.inf
        probs.append(("gemm", dict(M=M, N=N, K=K, ta=NO_TRANS, tb=NO_TRANS, A=A, B=B, C=C, alpha=1.5, beta=0.0)))
    m, n = 300, 500
    y = rng.uniform(-1, 1, m).astype(np.float32)
    y[::13] = np.nan
    probs.append(("gemv", dict(m=m, n=n, A=rng.uniform(-1, 1, (m, n)).astype(np.float32),
                               x=rng.uniform(-1, 1, n).astype(np.float32), y=y, alpha=0.75, beta=0.0)))
    ops, keep = _to_device(torch, native, probs)
    assert so.exoblas_batch_run(torch.cuda.current_stream().cuda_stream, ops.ctypes.data, len(ops)) == 0
    torch.cuda.synchronize()
    o = Oracle(np.float32)
    for kind, p in probs:
        got = p["out"].cpu().numpy()
        bad = ~np.isfinite(p["C"] if kind == "gemm" else p["y"])
        assert np.isnan(got[bad]).all(), f"batched {kind}: beta = 0 must keep NaN from 0 * NaN / Inf"
        if kind == "gemm":
            want = np.zeros_like(p["C"])
            o.gemm(NO_TRANS, NO_TRANS, p["M"], p["N"], p["K"], p["alpha"], p["A"], p["K"], p["B"], p["N"], want,
                   p["N"])
            mag = p["alpha"] * (np.abs(p["A"]).astype(np.float64) @ np.abs(p["B"]))
            K = p["K"]
        else:
            want = np.zeros(p["m"], np.float32)
            o.gemv(NO_TRANS, p["m"], p["n"], p["alpha"], 0.0, p["A"], p["n"], p["x"], 1, want, 1)
            mag = p["alpha"] * (np.abs(p["A"]).astype(np.float64) @ np.abs(p["x"]))
            K = p["n"]
        assert_bound(np.where(bad, 0, got), np.where(bad, 0, want), np.where(bad, 1, mag), K, np.float32,
                     what=f"batched {kind} with beta = 0")


@pytest.mark.parametrize("tile", [0, 1, 2])
def test_batch_every_tile(env, rng, tile):
    torch, so, native = env
    probs = [p for p in _mixed_problems(rng) if p[0] == "gemm"]
    ops, keep = _to_device(torch, native, probs)
    assert so.exoblas_batch_force_tile(tile) == 0
    try:
        assert so.exoblas_batch_run(torch.cuda.current_stream().cuda_stream, ops.ctypes.data, len(ops)) == 0
        torch.cuda.synchronize()
    finally:
        so.exoblas_batch_force_tile(-1)
    _check(probs)


def test_plan_graph_replay(env, rng):
    """A plan launch captured in a hipGraph, replayed 4 times with eager resets
    of the outputs between replays; each replay equals the eager launch."""
    torch, so, native = env
    probs = _mixed_problems(rng)
    ops, keep = _to_device(torch, native, probs)
    inits = [p["out"].clone() for _, p in probs]
    plan = ctypes.c_void_p()
    assert so.exoblas_batch_plan_create(ctypes.byref(plan), ops.ctypes.data, len(ops)) == 0
    try:
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            assert so.exoblas_batch_plan_launch(plan, s.cuda_stream) == 0
        s.synchronize()
        eager = [p["out"].clone() for _, p in probs]
        for (_, p), i0 in zip(probs, inits):
            p["out"].copy_(i0)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            assert so.exoblas_batch_plan_launch(plan, torch.cuda.current_stream().cuda_stream) == 0
        for rep in range(4):
            for (_, p), i0 in zip(probs, inits):
                p["out"].copy_(i0)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for (kind, p), e in zip(probs, eager):
                assert torch.equal(p["out"], e), f"replay {rep}: {kind} differs from the eager launch"
    finally:
        so.exoblas_batch_plan_destroy(plan)
    _check(probs)


def _small_stream(count, seed, rate):
    from exoblas_amd import serving
    s = serving.make_stream(count, rate=rate, seed=seed)
    # keep the test quick: cap sizes
    s.m[:] = np.minimum(s.m, 600)
    s.n[:] = np.minimum(s.n, 600)
    s.k[:] = np.minimum(s.k, 600)
    s.n[s.kind == 3] = s.m[s.kind == 3]
    return s


def test_server_alone_equals_mix(env):
    """Each request's result alone (a one-request stream) equals its result
    inside the mixed stream, bit for bit: a request's work items do not depend
    on what else is in the batch."""
    torch, so, native = env
    from exoblas_amd import serving
    s = _small_stream(400, seed=3, rate=50000.0)
    w = serving.Workload(s)
    stats, lat = serving.run(w, mode=0, streams=3, max_batch=64)
    assert stats["completed"] == len(s.kind) and stats["rejected"] == 0
    mixed = [w.output(i).clone() for i in range(len(s.kind))]
    pick = [int(np.flatnonzero(s.kind == k)[0]) for k in (1, 2, 3)] + [7, 123, 399]
    for i in pick:
        w.reset()
        torch.cuda.synchronize()
        serving.run(w, mode=0, streams=1, max_batch=64, ops=w.ops[i:i + 1], arrival_us=np.zeros(1))
        torch.cuda.synchronize()
        assert torch.equal(w.output(i), mixed[i]), f"request {i} (kind {s.kind[i]}) differs alone vs in the mix"


def test_server_batched_matches_per_call_and_oracle(env):
    torch, so, native = env
    from exoblas_amd import serving
    s = _small_stream(200, seed=4, rate=20000.0)
    w = serving.Workload(s)
    serving.run(w, mode=0, streams=2)
    batched = [w.output(i).clone() for i in range(len(s.kind))]
    w.reset()
    serving.run(w, mode=1, streams=2)
    o = Oracle(np.float32)
    gA, gB, vA, vx = (t.cpu().numpy() for t in (w.gA, w.gB, w.vA, w.vx))
    init = w.init.cpu().numpy()
    for i in range(0, len(s.kind), 7):
        per_call = w.output(i).cpu().numpy()
        b = batched[i].cpu().numpy()
        m, n, k = int(s.m[i]), int(s.n[i]), int(s.k[i])
        off = int(w.out_off[i])
        if s.kind[i] == 1:
            want = init[off:off + m * n].reshape(m, n).copy()
            o.mscal(m, n, 0.5, want, n)
            o.gemm(NO_TRANS, NO_TRANS, m, n, k, 1.0, np.ascontiguousarray(gA[:m, :k]), k,
                   np.ascontiguousarray(gB[:k, :n]), n, want, n)
            mag = 0.5 * np.abs(init[off:off + m * n].reshape(m, n)) + np.abs(gA[:m, :k]).astype(np.float64) @ np.abs(gB[:k, :n])
            K = k
        elif s.kind[i] == 2:
            want = init[off:off + m].copy()
            o.gemv(NO_TRANS, m, n, 1.0, 0.5, np.ascontiguousarray(vA[:m, :n]), n, vx[:n].copy(), 1, want, 1)
            mag = 0.5 * np.abs(init[off:off + m]) + np.abs(vA[:m, :n]).astype(np.float64) @ np.abs(vx[:n])
            K = n
        else:
            want = init[off:off + m].copy()
            T = w.tri[m].cpu().numpy()
            o.trsv(LOWER, NO_TRANS, NON_UNIT, m, want, 1, T, m)
            mag = np.abs(want) + 1
            K = m
        assert_bound(b, want, mag, K, np.float32, what=f"batched request {i}")
        assert_bound(per_call, want, mag, K, np.float32, what=f"per-call request {i}")


def test_server_budget(env):
    torch, so, native = env
    from exoblas_amd import serving
    s = _small_stream(120, seed=5, rate=100000.0)
    w = serving.Workload(s)
    nbytes = w.operand_bytes()
    budget = int(nbytes.max() * 3)
    stats, _ = serving.run(w, mode=0, budget=budget, streams=4)
    assert stats["completed"] == len(s.kind)
    assert 0 < stats["peak_bytes_in_flight"] <= budget
    # a request that can never fit is rejected, the rest still complete
    big = int(nbytes.max())
    stats, lat = serving.run(w, mode=0, budget=big - 1, streams=2)
    assert stats["rejected"] == int((nbytes > big - 1).sum()) >= 1
    assert stats["completed"] == len(s.kind) - stats["rejected"]
    # the rejection is reported through the thread's sticky status
    from exoblas_amd import lib
    assert lib.last_status() == 6 and "budget" in lib.last_error()
    lib.clear_error()
