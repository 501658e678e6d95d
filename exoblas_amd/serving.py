"""Mixed-size request streams for the serving loop (include/exoblas_amd_batch.h).

``make_stream`` draws the request mix: 60 % sgemm NN with M, N, K each
log-uniform in [16, 1024], 25 % strsv lower / no-trans / non-unit with n
log-uniform in [32, 2048] on a diagonally dominant matrix (off-diagonal
U[-1,1]/n, diagonal n + U[0,1]), 15 % sgemv N with m, n log-uniform in
[64, 8192]; Poisson arrivals at `rate` requests per second; everything from
one seeded generator.  ``Workload`` puts every operand in device memory
(torch tensors), either shared (every request reads the same few matrices) or
pooled (each request its own window of large pools), every request its own
output, and builds the exoblas_batch_op array.  ``run`` replays the stream
through a server and returns its statistics.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import lib as L
from . import native as N

NO_TRANS, LOWER, NON_UNIT = 111, 122, 131


def loguniform_int(rng, lo, hi, size):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size))).astype(np.int64).clip(lo, hi)


@dataclass
class Stream:
    kind: np.ndarray      # OP_GEMM / OP_GEMV / OP_TRSV
    m: np.ndarray
    n: np.ndarray
    k: np.ndarray
    arrival_us: np.ndarray


def make_stream(count: int = 20000, rate: float = 20000.0, seed: int = 0) -> Stream:
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, count)
    kind = np.where(u < 0.60, N.OP_GEMM, np.where(u < 0.85, N.OP_TRSV, N.OP_GEMV)).astype(np.int32)
    m = np.zeros(count, np.int64)
    n = np.zeros(count, np.int64)
    k = np.zeros(count, np.int64)
    g = kind == N.OP_GEMM
    m[g] = loguniform_int(rng, 16, 1024, g.sum())
    n[g] = loguniform_int(rng, 16, 1024, g.sum())
    k[g] = loguniform_int(rng, 16, 1024, g.sum())
    t = kind == N.OP_TRSV
    m[t] = loguniform_int(rng, 32, 2048, t.sum())
    n[t] = m[t]
    v = kind == N.OP_GEMV
    m[v] = loguniform_int(rng, 64, 8192, v.sum())
    n[v] = loguniform_int(rng, 64, 8192, v.sum())
    gaps = rng.exponential(1e6 / rate, count)
    return Stream(kind, m, n, k, np.cumsum(gaps) - gaps[0])


class Workload:
    """Device-resident operands for a stream, and its op array.

    Two layouts.  Shared (``pool_bytes`` None): the read-only gemm / gemv
    inputs are windows at the start of two 1024^2 matrices, an 8192^2 matrix
    and one x, each trsv order gets its own matrix.  Pooled: every request
    reads its own window at a random 256-byte aligned offset of pools of
    ``pool_bytes`` each (gemm A and B with ld 1024, gemv A with ld 8192 and
    its x), and each trsv takes the leading n x n block (ld 2048) of one of
    ``tri_pool`` 2048^2 diagonally dominant matrices, so consecutive requests
    do not find each other's operands in the caches.  Either way every
    request writes its own output window."""

    def __init__(self, s: Stream, seed: int = 1, pool_bytes: Optional[int] = None, tri_pool: int = 24):
        import torch
        self.torch = torch
        self.s = s
        self.pooled = pool_bytes is not None
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        dev = "cuda"
        rng = np.random.default_rng(seed)

        def u(*shape):
            return torch.empty(shape, dtype=torch.float32, device=dev).uniform_(-1, 1, generator=g)

        def dd(nn, ld):
            T = u(nn, ld) / nn
            T.diagonal().copy_(nn + torch.empty(nn, device=dev).uniform_(0, 1, generator=g))
            return T
        count = len(s.kind)
        self.src = [None] * count  # per request: (tensor, offset, ld) of A and of B / x
        if not self.pooled:
            self.gA, self.gB = u(1024, 1024), u(1024, 1024)
            self.vA, self.vx = u(8192, 8192), u(8192)
            self.tri: Dict[int, "torch.Tensor"] = {}
            for nn in np.unique(s.m[s.kind == N.OP_TRSV]):
                self.tri[int(nn)] = dd(int(nn), int(nn))
            for i in range(count):
                m, n, k = int(s.m[i]), int(s.n[i]), int(s.k[i])
                if s.kind[i] == N.OP_GEMM:
                    self.src[i] = ((self.gA, 0, 1024), (self.gB, 0, 1024))
                elif s.kind[i] == N.OP_GEMV:
                    self.src[i] = ((self.vA, 0, 8192), (self.vx, 0, 1))
                else:
                    self.src[i] = ((self.tri[m], 0, m), None)
        else:
            pe = max(int(pool_bytes) // 4, 8192 * 8192)
            self.pA, self.pB = u(pe), u(pe)
            self.pV = u(((pe + 8191) // 8192) * 8192)
            self.pX = u(pe // 4)
            self.tri_pool = [dd(2048, 2048) for _ in range(tri_pool)]

            def off(total, need):
                return int(rng.integers(0, (total - need) // 64 + 1)) * 64
            for i in range(count):
                m, n, k = int(s.m[i]), int(s.n[i]), int(s.k[i])
                if s.kind[i] == N.OP_GEMM:
                    self.src[i] = ((self.pA, off(pe, m * 1024), 1024), (self.pB, off(pe, k * 1024), 1024))
                elif s.kind[i] == N.OP_GEMV:
                    self.src[i] = ((self.pV, off(self.pV.numel(), m * 8192), 8192), (self.pX, off(self.pX.numel(), n), 1))
                else:
                    self.src[i] = ((self.tri_pool[int(rng.integers(0, tri_pool))], 0, 2048), None)
        # outputs: one arena, each request its own window
        out = np.where(s.kind == N.OP_GEMM, s.m * s.n, s.m)
        self.out_off = np.concatenate([[0], np.cumsum(out + 64)])  # 256-byte separation
        self.out = torch.empty(int(self.out_off[-1]), dtype=torch.float32, device=dev)
        self.init = u(int(self.out_off[-1]))  # initial C / y / right-hand sides
        self.reset()
        self.ops = self._ops()

    def reset(self):
        self.out.copy_(self.init)

    def output(self, i):
        s = self.s
        o = int(self.out_off[i])
        if s.kind[i] == N.OP_GEMM:
            return self.out[o:o + int(s.m[i] * s.n[i])].view(int(s.m[i]), int(s.n[i]))
        return self.out[o:o + int(s.m[i])]

    def operands(self, i):
        """Request i's read-only operands as host arrays: (A, B) for gemm, (A,
        x) for gemv, (T,) for trsv, each contiguous."""
        s = self.s
        m, n, k = int(s.m[i]), int(s.n[i]), int(s.k[i])
        (ta, oa, la), bsrc = self.src[i]

        def mat(t, o, ld, r, c):
            flat = t.reshape(-1)
            return flat[o:o + (r - 1) * ld + c].cpu().numpy() if r > 0 else np.zeros(0, np.float32), ld

        def dense(t, o, ld, r, c):
            v, ld = mat(t, o, ld, r, c)
            return np.ascontiguousarray(np.lib.stride_tricks.as_strided(v, (r, c), (4 * ld, 4)))
        if s.kind[i] == N.OP_GEMM:
            tb, ob, lb = bsrc
            return dense(ta, oa, la, m, k), dense(tb, ob, lb, k, n)
        if s.kind[i] == N.OP_GEMV:
            tb, ob, _ = bsrc
            return dense(ta, oa, la, m, n), tb.reshape(-1)[ob:ob + n].cpu().numpy()
        return (dense(ta, oa, la, m, m),)

    def _ops(self) -> np.ndarray:
        s = self.s
        ops = np.zeros(len(s.kind), N.OP_DTYPE)
        base = self.out.data_ptr()
        ops["kind"] = s.kind
        ops["trans_a"] = NO_TRANS
        ops["m"], ops["n"], ops["k"] = s.m, s.n, s.k
        ops["c"] = base + 4 * self.out_off[:-1]
        ops["alpha"] = 1.0
        g = s.kind == N.OP_GEMM
        ops["trans_b"][g] = NO_TRANS
        ops["ldc"][g] = s.n[g]
        ops["beta"][g] = 0.5
        v = s.kind == N.OP_GEMV
        ops["ldc"][v] = 1
        ops["beta"][v] = 0.5
        t = s.kind == N.OP_TRSV
        ops["trans_b"][t] = LOWER
        ops["diag"][t] = NON_UNIT
        ops["ldc"][t] = 1
        for i in range(len(s.kind)):
            (ta, oa, la), bsrc = self.src[i]
            ops["a"][i] = ta.data_ptr() + 4 * oa
            ops["lda"][i] = la
            if bsrc is not None:
                tb, ob, lb = bsrc
                ops["b"][i] = tb.data_ptr() + 4 * ob
                ops["ldb"][i] = lb
        return ops

    def operand_bytes(self) -> np.ndarray:
        s = self.s
        return np.where(s.kind == N.OP_GEMM, 4 * (s.m * s.k + s.k * s.n + s.m * s.n),
                        np.where(s.kind == N.OP_GEMV, 4 * (s.m * s.n + s.n + s.m), 4 * (s.m * (s.m + 1) // 2 + s.m)))


def run(w: Workload, mode: int, budget: int = 64 << 30, streams: int = 4, max_batch: int = 256,
        ops: Optional[np.ndarray] = None, arrival_us: Optional[np.ndarray] = None):
    """Replays the stream (or `ops` at `arrival_us`) through a fresh server;
    returns (stats dict, per-request latencies in us)."""
    from . import autotune
    so = L.load()
    autotune.apply()  # the tuned gemm tile table, when one has been written
    ops = w.ops if ops is None else ops
    arr = w.s.arrival_us if arrival_us is None else arrival_us
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    ops = np.ascontiguousarray(ops)
    srv = ctypes.c_void_p()
    st = so.exoblas_server_create(ctypes.byref(srv), budget, streams, mode, max_batch)
    if st != 0:
        raise RuntimeError(L.last_error())
    try:
        st = so.exoblas_server_replay(srv, ops.ctypes.data, arr.ctypes.data, len(ops))
        if st != 0:
            raise RuntimeError(L.last_error())
        st = so.exoblas_server_drain(srv)
        stats = N.ServerStats()
        so.exoblas_server_stats_get(srv, ctypes.byref(stats))
        lat = np.zeros(len(ops))
        so.exoblas_server_latencies(srv, lat.ctypes.data, len(ops))
        if st != 0:
            raise RuntimeError(f"server reported a failure: {L.last_error()}")
    finally:
        so.exoblas_server_destroy(srv)
    d = {f: getattr(stats, f) for f, _ in N.ServerStats._fields_}
    d["req_per_s"] = d["completed"] / d["wall_s"] if d["wall_s"] > 0 else 0.0
    return d, lat
