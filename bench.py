"""Bandwidth-set benchmark of the exo_* BLAS on one MI355X.

Workload (one "step" = one pass over the set, every operand resident in HBM):
  saxpy, daxpy   n = 2^26, alpha = 0.75, x, y ~ U[-1,1]
  sdot, ddot     n = 2^26
  sgemv, dgemv   N and T, 8192 x 8192 row-major, lda = 8192, alpha = 0.5, beta = 1.5
  scopy, sswap   n = 2^26 (bit-exact data movement)
  srot           n = 2^26, c = 0.6, s = 0.8
Each kernel is called through its device-pointer C-ABI entry point
(exoblas_<k>_stride_1_async) on torch's current stream.  Operands rotate
between two buffer sets (> 512 MB apart in total) so the 256 MB Infinity
Cache cannot serve a kernel's inputs from the previous step.

value = sum of algorithmic bytes over the set / measured time per step (GB/s).
Algorithmic bytes per call: axpy 3ns, dot 2ns, gemv (MN + N + 2M)s, copy 2ns,
swap 4ns, rot 4ns (s = element size).  roofline = the kernel with the largest
share of the step, timed with HIP events around its launches on the same
stream, against the 8 TB/s HBM3E peak.  cpu_baseline = the oracle's loop nests
built with the reference's avx512 preset flags, one thread, timed on this
host over a bounded sample of the same set.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GFLOP/s or GB/s and %roofline per kernel; mixed-size BLAS req/s + p50/p99"


def kernel_set(n1, n2):
    """(name, prefix, kind, bytes) for the bandwidth set."""
    out = []
    for p, s in (("s", 4), ("d", 8)):
        out.append((f"{p}axpy", p, "axpy", 3 * n1 * s))
        out.append((f"{p}dot", p, "dot", 2 * n1 * s))
        out.append((f"{p}gemv_N", p, "gemv_n", (n2 * n2 + 3 * n2) * s))
        out.append((f"{p}gemv_T", p, "gemv_t", (n2 * n2 + 3 * n2) * s))
    out.append(("scopy", "s", "copy", 2 * n1 * 4))
    out.append(("sswap", "s", "swap", 4 * n1 * 4))
    out.append(("srot", "s", "rot", 4 * n1 * 4))
    return out


class Operands:
    def __init__(self, torch, n1, n2, seed):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)

        def u(shape, dt):
            t = torch.empty(shape, dtype=dt, device="cuda")
            t.uniform_(-1.0, 1.0, generator=g)
            return t
        self.v = {}
        for p, dt in (("s", torch.float32), ("d", torch.float64)):
            self.v[p] = dict(x=u(n1, dt), y=u(n1, dt), A=u((n2, n2), dt), gx=u(n2, dt), gy=u(n2, dt),
                             res=torch.zeros(1, dtype=dt, device="cuda"))


def make_launchers(so, L, ops, n1, n2, stream):
    c_f = {"s": ctypes.c_float, "d": ctypes.c_double}
    alpha = {p: ctypes.byref(c_f[p](0.75)) for p in "sd"}
    ga = {p: ctypes.byref(c_f[p](0.5)) for p in "sd"}
    gb = {p: ctypes.byref(c_f[p](1.5)) for p in "sd"}
    rc, rs = ctypes.byref(ctypes.c_float(0.6)), ctypes.byref(ctypes.c_float(0.8))

    def w1(t):
        return L.win1(t.data_ptr(), 1)

    def w2(t, ld):
        return L.win2(t.data_ptr(), ld)

    def f(name, p, kind):
        o = ops.v[p]
        if kind == "axpy":
            g = getattr(so, f"exoblas_{p}axpy_stride_1_async")
            return lambda: g(stream, n1, alpha[p], w1(o["x"]), w1(o["y"]))
        if kind == "dot":
            g = getattr(so, f"exoblas_{p}dot_stride_1_async")
            return lambda: g(stream, n1, w1(o["x"]), w1(o["y"]), ctypes.c_void_p(o["res"].data_ptr()))
        if kind in ("gemv_n", "gemv_t"):
            g = getattr(so, f"exoblas_{p}gemv_rm_stride_1_async")
            tr = 111 if kind == "gemv_n" else 112
            return lambda: g(stream, tr, n2, n2, ga[p], gb[p], w2(o["A"], n2), w2(o["A"], n2), w1(o["gx"]),
                             w1(o["gy"]))
        if kind == "copy":
            g = so.exoblas_scopy_stride_1_async
            return lambda: g(stream, n1, w1(o["x"]), w1(o["y"]))
        if kind == "swap":
            g = so.exoblas_sswap_stride_1_async
            return lambda: g(stream, n1, w1(o["x"]), w1(o["y"]))
        if kind == "rot":
            g = so.exoblas_srot_stride_1_async
            return lambda: g(stream, n1, w1(o["x"]), w1(o["y"]), rc, rs)
        raise ValueError(kind)
    return f


def cpu_baseline(n1, n2, budget_s=15.0):
    """The oracle's loop nests (avx512 preset build), one thread, on a bounded
    sample: every kernel of the set at its full size, repeated while the
    budget lasts (at least once)."""
    import numpy as np

    from oracle import Oracle, _host_has_avx512
    rng = np.random.default_rng(0)
    total_bytes = 0.0
    total_s = 0.0
    reps_done = 0
    data = {}
    for p, dt in (("s", np.float32), ("d", np.float64)):
        data[p] = dict(x=rng.uniform(-1, 1, n1).astype(dt), y=rng.uniform(-1, 1, n1).astype(dt),
                       A=rng.uniform(-1, 1, (n2, n2)).astype(dt), gx=rng.uniform(-1, 1, n2).astype(dt),
                       gy=rng.uniform(-1, 1, n2).astype(dt))
    ks = kernel_set(n1, n2)
    t_start = time.perf_counter()
    while True:
        for name, p, kind, nbytes in ks:
            o = Oracle(np.float32 if p == "s" else np.float64, kind="fast")
            d = data[p]
            t0 = time.perf_counter()
            if kind == "axpy":
                o.axpy(n1, 0.75, d["x"], 1, d["y"], 1)
            elif kind == "dot":
                o.dot(n1, d["x"], 1, d["y"], 1)
            elif kind == "gemv_n":
                o.gemv(111, n2, n2, 0.5, 1.5, d["A"], n2, d["gx"], 1, d["gy"], 1)
            elif kind == "gemv_t":
                o.gemv(112, n2, n2, 0.5, 1.5, d["A"], n2, d["gx"], 1, d["gy"], 1)
            elif kind == "copy":
                o.copy(n1, d["x"], 1, d["y"], 1)
            elif kind == "swap":
                o.swap(n1, d["x"], 1, d["y"], 1)
            elif kind == "rot":
                o.rot(n1, d["x"], 1, d["y"], 1, 0.6, 0.8)
            total_s += time.perf_counter() - t0
            total_bytes += nbytes
        reps_done += 1
        if time.perf_counter() - t_start > budget_s:
            break
    return {"value": round(total_bytes / total_s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps_done} pass(es) over the full bandwidth set ({'avx512' if _host_has_avx512() else 'avx2'}"
                      f" -O3 -ffast-math build of the oracle, 1 thread, host-resident operands)"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n1", type=int, default=1 << 26)
    ap.add_argument("--n2", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2

    import torch

    from exoblas_amd import lib as L
    so = L.load()
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    sets = [Operands(torch, a.n1, a.n2, seed) for seed in (0, 1)]
    ks = kernel_set(a.n1, a.n2)
    launch = [[make_launchers(so, L, s, a.n1, a.n2, sh)(name, p, kind) for name, p, kind, _ in ks] for s in sets]

    def step(i, evs=None):
        fs = launch[i % 2]
        for j, f in enumerate(fs):
            if evs is not None:
                evs[j][0].record(stream)
            rc = f()
            if evs is not None:
                evs[j][1].record(stream)
            if rc != 0:
                raise RuntimeError(f"{ks[j][0]} failed: {L.last_error()}")

    for i in range(a.warmup):
        step(i)
    torch.cuda.synchronize()
    evs = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in ks]
           for _ in range(a.steps)]
    t0 = time.perf_counter()
    for i in range(a.steps):
        step(i, evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per_k_ms = [sum(evs[i][j][0].elapsed_time(evs[i][j][1]) for i in range(a.steps)) / a.steps
                for j in range(len(ks))]
    step_ms = wall * 1e3 / a.steps
    total_bytes = sum(b for *_, b in ks)
    value = total_bytes / (step_ms * 1e-3) / 1e9
    kernels = {}
    for (name, p, kind, b), ms in zip(ks, per_k_ms):
        gbs = b / (ms * 1e-3) / 1e9
        kernels[name] = {"us": round(ms * 1e3, 2), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "bytes": b}
    dom = max(range(len(ks)), key=lambda j: per_k_ms[j])
    dname, _, _, dbytes = ks[dom]
    dach = dbytes / (per_k_ms[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dname, {}).get("bytes_per_launch")
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(step_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic U[-1,1], device-resident",
        "config": {"workload": f"bandwidth set: s/d axpy+dot n={a.n1}, s/d gemv N+T {a.n2}x{a.n2}, "
                               f"scopy/sswap/srot n={a.n1}", "n": a.n1, "gemv": [a.n2, a.n2],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": round(dach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dach / HBM_PEAK_GBS, 4), "traffic": traffic},
        "kernels": kernels,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.n1, a.n2, a.cpu_budget)
        out["cpu_baseline"]["cores_on_host"] = os.cpu_count()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
