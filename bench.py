"""Benchmark of the exo_* BLAS on one MI355X.

Default run: one JSON line whose headline is the bandwidth set, with every
other configuration under "configs".

Headline workload (one "step" = one pass over the set, every operand resident
in HBM):
  saxpy, daxpy   n = 2^26, alpha = 0.75, x, y ~ U[-1,1]
  sdot, ddot     n = 2^26
  sgemv, dgemv   N and T, 8192 x 8192 row-major, lda = 8192, alpha = 0.5, beta = 1.5
  scopy, sswap   n = 2^26 (bit-exact data movement)
  srot           n = 2^26, c = 0.6, s = 0.8
Each kernel is called through its device-pointer C-ABI entry point
(exoblas_<k>_stride_1_async) on torch's current stream.  Two operand sets
(> 512 MB each) alternate from kernel to kernel, so the 256 MB Infinity Cache
never holds what the next kernel reads.

value = sum of algorithmic bytes over the set / measured time per step (GB/s).
Algorithmic bytes per call: axpy 3ns, dot 2ns, gemv (MN + N + 2M)s, copy 2ns,
swap 4ns, rot 4ns (s = element size).  roofline = the kernel with the largest
share of the step, timed with HIP events around its launches on the same
stream, against the 8 TB/s HBM3E peak.  cpu_baseline = the oracle's loop nests
built with the reference's avx512 preset flags, one thread, timed on this
host over a bounded sample of the same set, with the same loop nests on every
usable host thread (OpenMP) and the host's vendor CBLAS beside it.

configs: config1 (sgemm 512^3 with beta), config3 (triangular and symmetric),
config4 (split precision and the mixed solve), config5 (serving), families
(the kernels the set leaves out); each with its own roofline object and CPU
baseline, each runnable alone with --workload <name>.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

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


def host_cpus():
    """What this process may use of the host: the affinity mask and the
    cgroup CPU quota (a container can see more CPUs than its quota lets it
    run; threads beyond the quota only contend)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:  # cgroup v2: "<quota> <period>" or "max <period>"
            q, per = f.read().split()
            if q != "max":
                quota = int(q) / int(per)
    except (OSError, ValueError):
        try:
            with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
                q = int(f.read())
            with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                per = int(f.read())
            if q > 0:
                quota = q / per
        except (OSError, ValueError):
            pass
    use = aff if quota is None else max(1, min(aff, int(quota)))
    return {"os_cpu_count": os.cpu_count(), "affinity": aff, "cgroup_quota_cpus": quota, "threads": use}


def set_port_threads(n):
    """Thread count of the all-cores oracle build (its OpenMP runtime)."""
    from oracle import load
    load("omp")
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(ctypes.c_int(n))


def cpu_rate(fn, budget_s, min_reps=1):
    """(seconds per call, calls) of fn() repeated while the budget lasts,
    after one untimed call."""
    fn()
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if reps >= min_reps and el > budget_s:
            return el / reps, reps


def _mkl():
    import ctypes.util
    path = next((q for q in ("/opt/conda/lib/libmkl_rt.so", ctypes.util.find_library("mkl_rt") or "")
                 if q and os.path.exists(q)), None)
    return (ctypes.CDLL(path), os.path.basename(path)) if path else (None, None)


def cpu_baseline(n1, n2, budget_s=15.0, kind="fast", threads=1):
    """The oracle's loop nests (avx512 preset build), one thread ("fast") or
    every usable host thread ("omp"), on a bounded sample: every kernel of
    the set at its full size, repeated while the budget lasts (at least
    once)."""
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
            o = Oracle(np.float32 if p == "s" else np.float64, kind=kind)
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
    return {"value": round(total_bytes / total_s / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{reps_done} pass(es) over the full bandwidth set ({'avx512' if _host_has_avx512() else 'avx2'}"
                      f" -O3 -ffast-math build of the oracle, {threads} thread(s)"
                      f"{' (OpenMP over independent outer loops)' if kind == 'omp' else ''}, host-resident operands)"}


def vendor_cblas(n1, n2, budget_s, threads):
    """The same set through the host's vendor CBLAS (MKL's libmkl_rt, as the
    reference's own benchmarks compare against), `threads` threads; None
    when no such library is present."""
    import ctypes.util

    import numpy as np
    path = next((q for q in ("/opt/conda/lib/libmkl_rt.so", ctypes.util.find_library("mkl_rt") or "")
                 if q and os.path.exists(q)), None)
    if path is None:
        return None
    mkl = ctypes.CDLL(path)
    mkl.MKL_Set_Num_Threads(ctypes.c_int(threads))
    rng = np.random.default_rng(0)
    P = ctypes.c_void_p
    total_bytes = total_s = 0.0
    reps = 0
    t_start = time.perf_counter()
    data = {}
    for p, dt, T in (("s", np.float32, ctypes.c_float), ("d", np.float64, ctypes.c_double)):
        data[p] = (dt, T, rng.uniform(-1, 1, n1).astype(dt), rng.uniform(-1, 1, n1).astype(dt),
                   rng.uniform(-1, 1, (n2, n2)).astype(dt), rng.uniform(-1, 1, n2).astype(dt),
                   rng.uniform(-1, 1, n2).astype(dt))
    ptr = lambda a: a.ctypes.data_as(P)
    first = True  # one untimed pass: library initialisation and page faults
    while True:
        for name, p, kind, nbytes in kernel_set(n1, n2):
            dt, T, x, y, A, gx, gy = data[p]
            f = lambda k: getattr(mkl, f"cblas_{p}{k}")
            t0 = time.perf_counter()
            if kind == "axpy":
                f("axpy")(ctypes.c_int(n1), T(0.75), ptr(x), 1, ptr(y), 1)
            elif kind == "dot":
                fd = f("dot")
                fd.restype = T
                fd(ctypes.c_int(n1), ptr(x), 1, ptr(y), 1)
            elif kind in ("gemv_n", "gemv_t"):
                f("gemv")(101, 111 if kind == "gemv_n" else 112, n2, n2, T(1.5), ptr(A), n2, ptr(gx), 1, T(0.5),
                          ptr(gy), 1)
            elif kind == "copy":
                f("copy")(ctypes.c_int(n1), ptr(x), 1, ptr(y), 1)
            elif kind == "swap":
                f("swap")(ctypes.c_int(n1), ptr(x), 1, ptr(y), 1)
            elif kind == "rot":
                f("rot")(ctypes.c_int(n1), ptr(x), 1, ptr(y), 1, T(0.6), T(0.8))
            if not first:
                total_s += time.perf_counter() - t0
                total_bytes += nbytes
        if first:
            first = False
            t_start = time.perf_counter()
            continue
        reps += 1
        if time.perf_counter() - t_start > budget_s:
            break
    return {"value": round(total_bytes / total_s / 1e9, 3), "unit": "GB/s", "threads": threads,
            "sample": f"{reps} pass(es) over the full bandwidth set, {os.path.basename(path)}"}


def _gbench_emit(a, groups):
    """Writes {kernel file: [records]} under --gbench-dir (Google Benchmark
    layout read by the reference's analytics_tools/graphing)."""
    if not a.gbench_dir:
        return
    from exoblas_amd import gbench
    for kernel, recs in groups.items():
        gbench.write(a.gbench_dir, "exoblas_amd", kernel, recs)


def bench_bandwidth(a, torch, so, L):
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    sets = [Operands(torch, a.n1, a.n2, seed) for seed in (0, 1)]
    ks = kernel_set(a.n1, a.n2)
    launch = [[make_launchers(so, L, s, a.n1, a.n2, sh)(name, p, kind) for name, p, kind, _ in ks] for s in sets]

    def step(i, evs=None, only=None):
        # kernel j of step i uses operand set (i + j) % 2: consecutive kernels
        # (and the same kernel in consecutive steps) never share operands, so
        # the 256 MB Infinity Cache holds nothing the next kernel reads
        for j in range(len(ks)):
            f = launch[(i + j) % 2][j]
            timed = evs is not None and (only is None or j == only)
            if timed:
                evs[j][0].record(stream)
            rc = f()
            if timed:
                evs[j][1].record(stream)
            if rc != 0:
                raise RuntimeError(f"{ks[j][0]} failed: {L.last_error()}")

    for i in range(a.warmup):
        step(i)
    torch.cuda.synchronize()
    # the timed steps carry one event pair, around the kernel with the most
    # bytes (the roofline's kernel); events between the other kernels would be
    # instrumentation, not work: the per-kernel times come from as many steps
    # again, each kernel between its own pair of events
    dom = max(range(len(ks)), key=lambda j: ks[j][3])
    dev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in ks]
           for _ in range(a.steps)]
    t0 = time.perf_counter()
    for i in range(a.steps):
        step(i, dev[i], only=dom)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dom_ms = sum(dev[i][dom][0].elapsed_time(dev[i][dom][1]) for i in range(a.steps)) / a.steps
    evs = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in ks]
           for _ in range(a.steps)]
    for i in range(a.steps):
        step(i, evs[i])
    torch.cuda.synchronize()

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
    if a.gbench_dir:
        from exoblas_amd import gbench
        groups = {}
        for (name, p, kind, b), ms in zip(ks, per_k_ms):
            if kind in ("gemv_n", "gemv_t"):
                groups.setdefault("gemv", []).append(
                    gbench.gemv_record(p, 111 if kind == "gemv_n" else 112, a.n2, a.n2, ms * 1e3, a.steps))
            else:
                f, rec = gbench.level1_records(p, kind, a.n1, ms * 1e3, a.steps)
                groups.setdefault(f, []).append(rec)
        _gbench_emit(a, groups)
    dname, _, _, dbytes = ks[dom]
    dach = dbytes / (dom_ms * 1e-3) / 1e9
    # HBM bytes per call of the dominant kernel from the FETCH_SIZE / WRITE_SIZE
    # counter runs (scaled by a plain copy of known size), at the default sizes
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_bandwidth_set.json")
    if os.path.exists(tpath) and a.n1 == 1 << 26 and a.n2 == 8192:
        with open(tpath) as f:
            traffic = json.load(f)["kernels"].get(dname, {}).get("traffic")
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
        "reference_copy": _runtime_copy(torch, stream, sets, a),
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_leg("bandwidth", a)
    return out


def _runtime_copy(torch, stream, sets, a):
    """The HIP runtime's own device-to-device copy (hipMemcpyAsync, which a
    contiguous torch copy_ issues) of the same n-element FP32 vectors as
    scopy, two sets taken in turn: the read-write stream ceiling this box
    gives a vendor kernel, beside scopy / sswap / srot."""
    xs = [st.v["s"]["x"] for st in sets]
    ys = [st.v["s"]["y"] for st in sets]
    for k in range(4):
        ys[k % 2].copy_(xs[k % 2])
    torch.cuda.synchronize()
    reps, tot = 10, 0.0
    for r in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ys[r % 2].copy_(xs[r % 2])
        e1.record(stream)
        torch.cuda.synchronize()
        tot += e0.elapsed_time(e1)
    ms = tot / reps
    b = 2 * a.n1 * 4
    gbs = b / (ms * 1e-3) / 1e9
    return {"kernel": "hipMemcpyAsync device-to-device (runtime blit kernel)", "us": round(ms * 1e3, 2),
            "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": b}


# ---------------------------------------------------------------- other workloads
# Each returns one JSON object with the same top-level keys; they measure the
# other configurations (level-3 products, triangular solves, split precision,
# serving) and are run with --workload <name>.

FP32_PEAK_TF, FP64_PEAK_TF, BF16_PEAK_TF = 157.3, 78.6, 2500.0


_SPIN = []


def _spin(torch, stream, us=30.0):
    """Enqueues about `us` microseconds of device spin on `stream` (cycles
    per microsecond calibrated once), so that the next call is enqueued
    before the device reaches it: the events around the call then bracket
    device time only, not the host's enqueue time (which matters for calls of
    a few microseconds)."""
    if not _SPIN:
        with torch.cuda.stream(stream):
            torch.cuda._sleep(1000)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            torch.cuda._sleep(200000)
            e1.record(stream)
        torch.cuda.synchronize()
        _SPIN.append(200000 / max(e0.elapsed_time(e1) * 1e3, 1e-3))  # cycles per us
    with torch.cuda.stream(stream):
        torch.cuda._sleep(max(1000, int(_SPIN[0] * us)))


def _timed(torch, stream, fn, reps, before=None):
    """Average ms of fn() over reps launches, HIP events on `stream`, each
    call's events enqueued behind a device spin (_spin)."""
    ts = []
    for _ in range(reps):
        if before is not None:
            before()
        _spin(torch, stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rc = fn()
        e1.record(stream)
        if rc not in (0, None):
            from exoblas_amd import lib as L
            raise RuntimeError(L.last_error())
        ts.append((e0, e1))
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ts) / reps


def _timed_batch(torch, stream, fn, reps, per=20):
    """Average device ms per call of fn() for calls of a few microseconds:
    `per` calls back to back between one event pair, behind a device spin
    long enough for the host to enqueue them all, so neither the host's
    enqueue time nor the events' own cost is in the figure; inter-kernel
    gaps of a stream are."""
    tot = 0.0
    for _ in range(reps):
        _spin(torch, stream, 40.0 * per)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _k in range(per):
            rc = fn()
            if rc not in (0, None):
                from exoblas_amd import lib as L
                raise RuntimeError(L.last_error())
        e1.record(stream)
        torch.cuda.synchronize()
        tot += e0.elapsed_time(e1) / per
    return tot / reps


def _line(metric, value, unit, a, ms, dtype, workload, roof, kernels, higher=True, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(ms, 4), "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic U[-1,1], device-resident",
           "config": {"workload": workload, "parallelism": "single GPU"}, "roofline": roof, "kernels": kernels}
    out.update(extra)
    return out


def bench_level3(a, torch, so, L):
    """sgemm / dgemm NN 8192^3, s/d syrk lower-N N=4096 K=1024, s/d symm
    left-lower 4096^2, sgemm NN 512^3 (alpha 1.5, beta applied by the
    separate mscal pass and not timed here)."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    kernels = {}
    total_flop, total_ms = 0.0, 0.0

    def u(shape, dt):
        return torch.empty(shape, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)
    cases = []
    for p, dt, peak in (("s", torch.float32, FP32_PEAK_TF), ("d", torch.float64, FP64_PEAK_TF)):
        T = ctypes.c_float if p == "s" else ctypes.c_double
        al = ctypes.byref(T(1.5))
        n = 8192
        A, B, C = u((n, n), dt), u((n, n), dt), u((n, n), dt)
        f = getattr(so, f"exoblas_{p}gemm_stride_1_async")
        cases.append((f"{p}gemm_NN_8192", peak, 2.0 * n ** 3,
                      (lambda f=f, A=A, B=B, C=C, al=al, n=n: f(sh, 111, 111, n, n, n, al, L.win2(A.data_ptr(), n),
                                                                 L.win2(B.data_ptr(), n), L.win2(A.data_ptr(), n),
                                                                 L.win2(B.data_ptr(), n), L.win2(C.data_ptr(), n))),
                      (A, B, C)))
        N, K = 4096, 1024
        As, Cs = u((N, K), dt), u((N, N), dt)
        f = getattr(so, f"exoblas_{p}syrk_rm_stride_1_async")
        w = L.win2(As.data_ptr(), K)
        cases.append((f"{p}syrk_LN_4096x1024", peak, float(K) * N * (N + 1),
                      (lambda f=f, w=w, Cs=Cs, al=al: f(sh, 122, 111, N, K, al, w, w, w, w, L.win2(Cs.data_ptr(), N))),
                      (As, Cs)))
        M = 4096
        Sa, Bm, Cm = u((M, M), dt), u((M, M), dt), u((M, M), dt)
        f = getattr(so, f"exoblas_{p}symm_rm_stride_1_async")
        cases.append((f"{p}symm_LL_4096", peak, 2.0 * M * M * M,
                      (lambda f=f, Sa=Sa, Bm=Bm, Cm=Cm, al=al: f(sh, 141, 122, M, M, al, L.win2(Sa.data_ptr(), M),
                                                                  L.win2(Sa.data_ptr(), M), L.win2(Bm.data_ptr(), M),
                                                                  L.win2(Cm.data_ptr(), M))),
                      (Sa, Bm, Cm)))
    n = 512
    A5, B5, C5 = u((n, n), torch.float32), u((n, n), torch.float32), u((n, n), torch.float32)
    cases.append(("sgemm_NN_512", FP32_PEAK_TF, 2.0 * n ** 3,
                  (lambda n=n: so.exoblas_sgemm_stride_1_async(sh, 111, 111, n, n, n, ctypes.byref(ctypes.c_float(1.5)),
                                                           L.win2(A5.data_ptr(), n), L.win2(B5.data_ptr(), n),
                                                           L.win2(A5.data_ptr(), n), L.win2(B5.data_ptr(), n),
                                                           L.win2(C5.data_ptr(), n))), (A5, B5, C5)))
    gb_groups = {}
    for name, peak, flop, fn, _ in cases:
        for _ in range(max(1, a.warmup // 2)):
            fn()
        reps = 20 if flop < 1e10 else max(2, min(a.steps, 5))
        ms = _timed(torch, stream, fn, reps)
        tf = flop / (ms * 1e-3) / 1e12
        kernels[name] = {"ms": round(ms, 4), "TFLOP/s": round(tf, 2), "frac": round(tf / peak, 4),
                         "peak": peak, "flop": flop}
        total_flop += flop
        total_ms += ms
        if a.gbench_dir:
            from exoblas_amd import gbench
            p, op = name[0], name[1:5]
            if op == "gemm":
                gn = int(name.rsplit("_", 1)[1])
                rec = gbench.gemm_record(p, 111, 111, gn, gn, gn, ms * 1e3, reps)
            elif op == "syrk":
                rec = gbench.syrk_record(p, 122, 111, 4096, 1024, ms * 1e3, reps)
            else:
                rec = gbench.symm_record(p, 141, 122, 4096, 4096, ms * 1e3, reps)
            gb_groups.setdefault(op, []).append(rec)
    _gbench_emit(a, gb_groups)
    dom = "sgemm_NN_8192"
    k = kernels[dom]
    roof = {"bound": "mfma", "kernel": dom, "achieved": k["TFLOP/s"], "peak": FP32_PEAK_TF, "unit": "TFLOP/s",
            "frac": k["frac"], "traffic": None}
    return _line("TFLOP/s per level-3 kernel (MFMA roofline)", round(total_flop / (total_ms * 1e-3) / 1e12, 2),
                 "TFLOP/s", a, total_ms, "f32+f64",
                 "level-3: s/d gemm NN 8192^3, s/d syrk L-N 4096x1024, s/d symm L-L 4096^2, sgemm NN 512^3",
                 roof, kernels)


def bench_trsv(a, torch, so, L):
    """s/d trsv lower, no-trans, non-unit at n = 4096 and 8192 on the
    diagonally dominant matrices (off-diagonal U[-1,1]/n, diagonal
    n + U[0,1]), b ~ U[-1,1]; x reset before every timed solve."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    kernels = {}
    gb = []
    for p, dt, s in (("s", torch.float32, 4), ("d", torch.float64, 8)):
        for n in (4096, 8192):
            A = torch.empty((n, n), dtype=dt, device="cuda").uniform_(-1, 1, generator=g) / n
            A.diagonal().copy_(n + torch.empty(n, dtype=dt, device="cuda").uniform_(0, 1, generator=g))
            b = torch.empty(n, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)
            x = b.clone()
            f = getattr(so, f"exoblas_{p}trsv_rm_stride_1_async")
            fn = (lambda f=f, A=A, x=x, n=n: f(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), L.win2(A.data_ptr(), n)))
            fn()
            ms = _timed(torch, stream, fn, 20, before=lambda x=x, b=b: x.copy_(b))
            nbytes = (n * (n + 1) // 2 + 2 * n) * s
            gbs = nbytes / (ms * 1e-3) / 1e9
            kernels[f"{p}trsv_LNN_{n}"] = {"us": round(ms * 1e3, 2), "GB/s": round(gbs, 1),
                                           "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": nbytes}
            if a.gbench_dir:
                from exoblas_amd import gbench
                gb.append(gbench.trsv_record(p, 122, 111, 131, n, ms * 1e3, 20))
    _gbench_emit(a, {"trsv": gb} if gb else {})
    k = kernels["strsv_LNN_4096"]
    roof = {"bound": "hbm", "kernel": "strsv_LNN_4096", "achieved": k["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": k["frac"], "traffic": None}
    tot_b = sum(v["bytes"] for v in kernels.values())
    tot_ms = sum(v["us"] for v in kernels.values()) / 1e3
    return _line("GB/s per triangular solve (HBM roofline; latency-bound chain)", round(tot_b / tot_ms / 1e6, 1),
                 "GB/s", a, tot_ms, "f32+f64", "trsv: s/d lower no-trans non-unit, n = 4096 and 8192", roof, kernels)


def bench_extended(a, torch, so, L):
    """The reference's other library targets: s/d syr2k lower-N N=4096 K=1024
    (2KN(N+1) flop), strmm Left-Lower-NoTrans-NonUnit 4096 x 4096 (M(M+1)N
    flop; B restored before each timed call, outside the timed region), and
    the banded level-2 kernels on n = 2^22 rows (s/d gbmv kl = ku = 16, sbmv
    upper k = 16, tbmv upper no-trans k = 16; band + vector bytes) and tbsv
    (upper no-trans, k = 16, n = 2^16, and k = 1024, n = 4096; latency-bound
    chains)."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    kernels = {}

    def u(shape, dt):
        return torch.empty(shape, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)

    def run(name, fn, work, unit, peak, reps, before=None):
        for _ in range(2):
            if before is not None:
                before()
            fn()
        ms = _timed(torch, stream, fn, reps, before=before)
        rate = work / (ms * 1e-3) / (1e12 if unit == "TFLOP/s" else 1e9)
        kernels[name] = {"us": round(ms * 1e3, 2), unit: round(rate, 2 if unit == "TFLOP/s" else 1),
                         "frac": round(rate / peak, 4), ("flop" if unit == "TFLOP/s" else "bytes"): work}

    for p, dt, s, peak in (("s", torch.float32, 4, FP32_PEAK_TF), ("d", torch.float64, 8, FP64_PEAK_TF)):
        T = ctypes.c_float if p == "s" else ctypes.c_double
        al, be = ctypes.byref(T(1.5)), ctypes.byref(T(0.5))
        N, K = 4096, 1024
        A, B, Cm = u((N, K), dt), u((N, K), dt), u((N, N), dt)
        f = getattr(so, f"exoblas_{p}syr2k_rm_stride_1_async")
        wa, wb = L.win2(A.data_ptr(), K), L.win2(B.data_ptr(), K)
        run(f"{p}syr2k_LN_4096x1024",
            lambda f=f, wa=wa, wb=wb, Cm=Cm, al=al: f(sh, 122, 111, N, K, al, wa, wa, wb, wb, wa, wa, wb, wb,
                                                      L.win2(Cm.data_ptr(), N)),
            2.0 * K * N * (N + 1), "TFLOP/s", peak, 5)
        n, k = 1 << 22, 16
        Ab = u((n, 2 * k + 1), dt)
        x, y = u((n,), dt), u((n,), dt)
        f = getattr(so, f"exoblas_{p}gbmv_row_major_NonTrans_stride_1_async")
        run(f"{p}gbmv_{n}_kl{k}_ku{k}",
            lambda f=f, Ab=Ab, x=x, y=y: f(sh, n, n, k, k, al, be, L.win2(Ab.data_ptr(), 2 * k + 1),
                                           L.win1(x.data_ptr(), 1), L.win1(y.data_ptr(), 1)),
            (n * (2 * k + 1) + n + 2 * n) * s, "GB/s", HBM_PEAK_GBS, 20)
        As = u((n, k + 1), dt)
        f = getattr(so, f"exoblas_{p}sbmv_row_major_Upper_stride_1_async")
        run(f"{p}sbmv_U_{n}_k{k}",
            lambda f=f, As=As, x=x, y=y: f(sh, n, k, al, L.win2(As.data_ptr(), k + 1), L.win1(x.data_ptr(), 1),
                                           L.win1(y.data_ptr(), 1)),
            (n * (k + 1) + n + 2 * n) * s, "GB/s", HBM_PEAK_GBS, 20)
        At = As.clone()
        At[:, 0] = 1.0
        xt = x.clone()
        f = getattr(so, f"exoblas_{p}tbmv_row_major_Upper_NonTrans_stride_1_async")
        run(f"{p}tbmv_UN_{n}_k{k}",
            lambda f=f, At=At, xt=xt: f(sh, n, k, L.win1(xt.data_ptr(), 1), L.win2(At.data_ptr(), k + 1), 0),
            (n * (k + 1) + 2 * n) * s, "GB/s", HBM_PEAK_GBS, 20, before=lambda xt=xt, x=x: xt.copy_(x))
        ns = 1 << 16
        Ad = u((ns, k + 1), dt)
        Ad[:, 0] = k + 1 + torch.empty(ns, dtype=dt, device="cuda").uniform_(0, 1, generator=g)
        b, xs = u((ns,), dt), torch.empty(ns, dtype=dt, device="cuda")
        f = getattr(so, f"exoblas_{p}tbsv_row_major_Upper_NonTrans_stride_1_async")
        run(f"{p}tbsv_UN_{ns}_k{k}",
            lambda f=f, Ad=Ad, xs=xs: f(sh, ns, k, L.win1(xs.data_ptr(), 1), L.win2(Ad.data_ptr(), k + 1), 0),
            (ns * (k + 1) + 2 * ns) * s, "GB/s", HBM_PEAK_GBS, 5, before=lambda xs=xs, b=b: xs.copy_(b))
        nw, kw = 4096, 1024  # the reference benchmark's largest tbsv shape (N = 4096, K = N / 4)
        Aw = u((nw, kw + 1), dt)
        Aw[:, 0] = kw + 1 + torch.empty(nw, dtype=dt, device="cuda").uniform_(0, 1, generator=g)
        bw, xw = u((nw,), dt), torch.empty(nw, dtype=dt, device="cuda")
        run(f"{p}tbsv_UN_{nw}_k{kw}",
            lambda f=f, Aw=Aw, xw=xw: f(sh, nw, kw, L.win1(xw.data_ptr(), 1), L.win2(Aw.data_ptr(), kw + 1), 0),
            (nw * (kw + 1) + 2 * nw) * s, "GB/s", HBM_PEAK_GBS, 5, before=lambda xw=xw, bw=bw: xw.copy_(bw))
    M = 4096
    At, B0 = u((M, M), torch.float32), u((M, M), torch.float32)
    Bw = B0.clone()
    run("strmm_LLNN_4096",
        lambda: so.exoblas_strmm_row_major_Left_Lower_NonTrans_async(sh, M, M, ctypes.byref(ctypes.c_float(1.0)),
                                                                     L.win2(At.data_ptr(), M),
                                                                     L.win2(Bw.data_ptr(), M), 0),
        float(M) * (M + 1) * M, "TFLOP/s", FP32_PEAK_TF, 5, before=lambda: Bw.copy_(B0))
    if a.gbench_dir:
        from exoblas_amd import gbench
        groups = {"syr2k": [gbench.syr2k_record(p, 122, 111, 4096, 1024, kernels[f"{p}syr2k_LN_4096x1024"]["us"], 5,
                                                alpha=1.5, beta=1) for p in "sd"],
                  "trmm": [gbench.strmm_record(122, 111, 131, M, M, 1, kernels["strmm_LLNN_4096"]["us"], 5)]}
        _gbench_emit(a, groups)
    k = kernels["ssyr2k_LN_4096x1024"]
    roof = {"bound": "mfma", "kernel": "ssyr2k_LN_4096x1024", "achieved": k["TFLOP/s"], "peak": FP32_PEAK_TF,
            "unit": "TFLOP/s", "frac": k["frac"], "traffic": None}
    tot_ms = sum(v["us"] for v in kernels.values()) / 1e3
    return _line("TFLOP/s (syr2k, trmm) and GB/s (banded) per kernel", k["TFLOP/s"], "TFLOP/s", a, tot_ms,
                 "f32+f64", "extended: s/d syr2k L-N 4096x1024, strmm L-L-N-N 4096^2, s/d gbmv/sbmv/tbmv n=2^22 "
                 "k=16, s/d tbsv n=2^16 k=16 and n=4096 k=1024", roof, kernels)


def bench_precision(a, torch, so, L):
    """sgemm NN 8192^3 and ssyrk L-N 8192 x 2048: plain FP32 MFMA vs split
    BF16 (3 and 6 products); max error against float64 on U[-1,1] and on rows
    scaled by 2^U[-8,8], as max |C - C64| / (|A| |B|) (normwise per element)
    and max relative.  strsv L-N-N n = 8192: plain FP32 solve vs the mixed
    solve (BF16 / FP32 updates, FP64 diagonal blocks, one refinement step),
    error against the FP64 solution."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    n = a.n_gemm
    kernels = {}
    one = ctypes.byref(ctypes.c_float(1.0))
    for scaled in (False, True):
        A = torch.empty((n, n), device="cuda").uniform_(-1, 1, generator=g)
        B = torch.empty((n, n), device="cuda").uniform_(-1, 1, generator=g)
        if scaled:
            A *= torch.exp2(torch.empty((n, 1), device="cuda").uniform_(-8, 8, generator=g))
            B *= torch.exp2(torch.empty((n, 1), device="cuda").uniform_(-8, 8, generator=g))
        truth = A.double() @ B.double()
        mag = A.double().abs() @ B.double().abs()
        C = torch.zeros((n, n), device="cuda")
        wa, wb, wc = L.win2(A.data_ptr(), n), L.win2(B.data_ptr(), n), L.win2(C.data_ptr(), n)
        forms = {"plain": lambda: so.exoblas_sgemm_stride_1_async(sh, 111, 111, n, n, n, one, wa, wb, wa, wb, wc),
                 "split3": lambda: so.exoblas_sgemm_split_async(sh, 111, 111, n, n, n, one, wa, wb, wa, wb, wc, 3),
                 "split4": lambda: so.exoblas_sgemm_split_async(sh, 111, 111, n, n, n, one, wa, wb, wa, wb, wc, 4),
                 "split6": lambda: so.exoblas_sgemm_split_async(sh, 111, 111, n, n, n, one, wa, wb, wa, wb, wc, 6)}
        for name, fn in forms.items():
            C.zero_()
            fn()
            torch.cuda.synchronize()
            err = (C.double() - truth).abs()
            nw = (err / mag).max().item()
            rel = (err / truth.abs().clamp_min(1e-300)).max().item()
            key = f"sgemm_{name}_{'scaled' if scaled else 'uniform'}"
            ms = _timed(torch, stream, fn, 3) if not scaled else None
            kernels[key] = {"max_err_normwise": nw, "max_rel_err": rel}
            if ms is not None:
                tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
                peak = FP32_PEAK_TF if name == "plain" else BF16_PEAK_TF / int(name[5:])
                kernels[key].update({"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "frac": round(tf / peak, 4),
                                     "peak_equiv": round(peak, 1)})
        del truth, mag, A, B, C
        torch.cuda.empty_cache()
    # ssyrk lower-N, N = 8192, K = 2048: plain vs split, error on the stored triangle
    N, K = a.n_gemm, a.n_gemm // 4
    for scaled in (False, True):
        A = torch.empty((N, K), device="cuda").uniform_(-1, 1, generator=g)
        if scaled:
            A *= torch.exp2(torch.empty((N, 1), device="cuda").uniform_(-8, 8, generator=g))
        A64 = A.double()
        tri = torch.ones((N, N), dtype=torch.bool, device="cuda").tril_()
        truth = A64 @ A64.T
        mag = A64.abs() @ A64.abs().T
        C = torch.zeros((N, N), device="cuda")
        wa, wc = L.win2(A.data_ptr(), K), L.win2(C.data_ptr(), N)
        forms = {"plain": lambda: so.exoblas_ssyrk_rm_stride_1_async(sh, 122, 111, N, K, one, wa, wa, wa, wa, wc),
                 "split3": lambda: so.exoblas_ssyrk_split_async(sh, 122, 111, N, K, one, wa, wa, wc, 3),
                 "split4": lambda: so.exoblas_ssyrk_split_async(sh, 122, 111, N, K, one, wa, wa, wc, 4),
                 "split6": lambda: so.exoblas_ssyrk_split_async(sh, 122, 111, N, K, one, wa, wa, wc, 6)}
        for name, fn in forms.items():
            C.zero_()
            fn()
            torch.cuda.synchronize()
            err = (C.double() - truth).abs()
            nw = (err / mag)[tri].max().item()
            key = f"ssyrk_{name}_{'scaled' if scaled else 'uniform'}"
            kernels[key] = {"max_err_normwise": nw}
            if not scaled:
                ms = _timed(torch, stream, fn, 3)
                flop = float(K) * N * (N + 1)
                tf = flop / (ms * 1e-3) / 1e12
                peak = FP32_PEAK_TF if name == "plain" else BF16_PEAK_TF / int(name[5:])
                kernels[key].update({"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "frac": round(tf / peak, 4),
                                     "peak_equiv": round(peak, 1)})
        del A, A64, tri, truth, mag, C
        torch.cuda.empty_cache()
    # strsv lower-N-nonunit, n = 8192 (diagonally dominant, as the trsv set):
    # plain FP32 solve vs BF16 / FP32 off-diagonal updates + FP64 diagonal
    # blocks + one refinement step; error against the FP64 solution.
    n = a.n_gemm
    A = torch.empty((n, n), device="cuda").uniform_(-1, 1, generator=g) / n
    A.diagonal().copy_(n + torch.empty(n, device="cuda").uniform_(0, 1, generator=g))
    b = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    x64 = torch.linalg.solve_triangular(A.double().tril(), b.double()[:, None], upper=False)[:, 0]
    x = b.clone()
    wA = L.win2(A.data_ptr(), n)
    packed = torch.empty(n * n, dtype=torch.bfloat16, device="cuda")  # the BF16 copy, made once
    assert so.exoblas_strsv_mixed_pack_async(sh, n, wA, packed.data_ptr()) == 0, L.last_error()
    forms = {"plain": lambda: so.exoblas_strsv_rm_stride_1_async(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA),
             "mixed_bf16": lambda: so.exoblas_strsv_mixed_async(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA,
                                                                16, 1),
             "mixed_bf16_packed": lambda: so.exoblas_strsv_mixed_packed_async(
                 sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA, packed.data_ptr(), 1),
             "mixed_fp32": lambda: so.exoblas_strsv_mixed_async(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA,
                                                                32, 1)}
    for name, fn in forms.items():
        x.copy_(b)
        fn()
        torch.cuda.synchronize()
        err = (x.double() - x64).abs()
        ms = _timed(torch, stream, fn, 10, before=lambda: x.copy_(b))
        nbytes = (n * (n + 1) // 2 + 2 * n) * 4
        kernels[f"strsv_{name}_{n}"] = {"max_err_normwise": (err.max() / x64.abs().max()).item(),
                                         "max_rel_err": (err / x64.abs().clamp_min(1e-300)).max().item(),
                                         "us": round(ms * 1e3, 2), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    del A, b, x, x64, packed
    torch.cuda.empty_cache()
    # the headline form: 4 products (the fastest whose error is at most plain
    # FP32's on both sets)
    sp = kernels["sgemm_plain_uniform"]["ms"] / kernels["sgemm_split4_uniform"]["ms"]
    k4 = kernels["sgemm_split4_uniform"]
    roof = {"bound": "mfma", "kernel": "sgemm_split4", "achieved": k4["TFLOP/s"], "peak": k4["peak_equiv"],
            "unit": "TFLOP/s (FP32-equivalent; BF16 peak / products)", "frac": k4["frac"], "traffic": None}
    err_ok = {s_: kernels[f"sgemm_split4_{s_}"]["max_err_normwise"] <= kernels[f"sgemm_plain_{s_}"]["max_err_normwise"]
              for s_ in ("uniform", "scaled")}
    return _line("split-BF16 sgemm speedup over plain FP32 MFMA sgemm", round(sp, 3), "x", a,
                 k4["ms"], "f32 (bf16 x3 / x4 / x6 products, f32 accumulate)",
                 f"sgemm NN {a.n_gemm}^3 and ssyrk L-N {a.n_gemm}x{a.n_gemm // 4}, uniform and row-scaled sets; "
                 f"strsv L-N-N n={a.n_gemm} plain vs mixed", roof, kernels, split4_error_le_plain=err_ok)


def bench_serving(a, torch, so, L):
    """The mixed request stream (seeded), batched persistent kernels vs one
    launch per call, same requests, same device-resident operands, 64 GB
    admission budget.  Two arrival patterns: the Poisson stream at --rate
    (latency under load) and the same requests arriving all at once
    (capacity: req/s when the queue never runs dry).  value = batched capacity."""
    from exoblas_amd import serving
    s = serving.make_stream(a.requests, rate=a.rate, seed=0)
    w = serving.Workload(s)
    res = {}
    for pattern, arrival in (("poisson", None), ("burst", np.zeros(a.requests))):
        for mode, name in ((1, "per_call"), (0, "batched")):
            w.reset()
            torch.cuda.synchronize()
            st, lat = serving.run(w, mode=mode, budget=64 << 30, streams=4, max_batch=a.max_batch,
                                  arrival_us=arrival)
            res[f"{pattern}_{name}"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}
    cap = res["burst_batched"]["req_per_s"]
    ratio = cap / max(res["burst_per_call"]["req_per_s"], 1e-9)
    roof = {"bound": "mixed", "kernel": "batch_kernel", "achieved": None, "peak": None, "unit": None, "frac": None,
            "traffic": None}
    return _line("mixed-size BLAS req/s (batched, capacity)", round(cap, 1), "req/s", a,
                 res["burst_batched"]["wall_s"] * 1e3, "f32",
                 f"serving: {a.requests} requests, 60% sgemm / 25% strsv / 15% sgemv; Poisson {a.rate:.0f}/s "
                 f"and all-at-once arrivals", roof, res, speedup_vs_per_call=round(ratio, 3),
                 p50_us=res["poisson_batched"]["p50_us"], p99_us=res["poisson_batched"]["p99_us"],
                 p50_us_per_call=res["poisson_per_call"]["p50_us"], p99_us_per_call=res["poisson_per_call"]["p99_us"])


def bench_hostpath(a, torch, so, L):
    """The drop-in path as the reference's harness drives it: synchronous
    exo_* calls on host (numpy) buffers, so every call stages its windows
    over PCIe and returns with the results in host memory.  Rates here are
    PCIe-inclusive (algorithmic bytes or flops / wall time per call); the
    oracle's avx512 build on one host thread is timed on the same buffers.
    sgemm 512^3 is config 1, with beta applied by the separate mscal call as
    the wrapper does (test/level3/gemm/exo_gemm_wrapper.h:18-20)."""
    import numpy as np

    from oracle import Oracle
    rng = np.random.default_rng(0)
    f32 = np.float32
    n1, n2, n3, nt = 1 << 26, 8192, 512, 4096
    x, y = rng.uniform(-1, 1, n1).astype(f32), rng.uniform(-1, 1, n1).astype(f32)
    A2, gx, gy = rng.uniform(-1, 1, (n2, n2)).astype(f32), rng.uniform(-1, 1, n2).astype(f32), np.zeros(n2, f32)
    A3, B3, C3 = (rng.uniform(-1, 1, (n3, n3)).astype(f32) for _ in range(3))
    T = (rng.uniform(-1, 1, (nt, nt)) / nt).astype(f32)
    T[np.arange(nt), np.arange(nt)] = (nt + rng.uniform(0, 1, nt)).astype(f32)
    b = rng.uniform(-1, 1, nt).astype(f32)
    xt = b.copy()
    res = np.zeros(1, f32)
    P = lambda arr: arr.ctypes.data
    fl = lambda v: ctypes.byref(ctypes.c_float(v))
    cases = {
        "saxpy_2^26": (3 * n1 * 4, "B",
                       lambda: so.exo_saxpy_stride_1(None, n1, fl(0.75), L.win1(P(x), 1), L.win1(P(y), 1)),
                       lambda o: o.axpy(n1, 0.75, x, 1, y, 1)),
        "sdot_2^26": (2 * n1 * 4, "B",
                      lambda: so.exo_sdot_stride_1(None, n1, L.win1(P(x), 1), L.win1(P(y), 1),
                                                   ctypes.c_void_p(P(res))),
                      lambda o: o.dot(n1, x, 1, y, 1)),
        "sgemv_N_8192": ((n2 * n2 + 3 * n2) * 4, "B",
                         lambda: so.exo_sgemv_rm_stride_1(None, 111, n2, n2, fl(1.5), fl(0.5), L.win2(P(A2), n2),
                                                          L.win2(P(A2), n2), L.win1(P(gx), 1), L.win1(P(gy), 1)),
                         lambda o: o.gemv(111, n2, n2, 1.5, 0.5, A2, n2, gx, 1, gy, 1)),
        "sgemm_NN_512": (2.0 * n3 ** 3, "flop",
                         lambda: (so.exo_smscal_rm_stride_1(None, n3, n3, fl(0.5), L.win2(P(C3), n3)),
                                  so.exo_sgemm_stride_1(None, 111, 111, n3, n3, n3, fl(1.5), L.win2(P(A3), n3),
                                                        L.win2(P(B3), n3), L.win2(P(A3), n3), L.win2(P(B3), n3),
                                                        L.win2(P(C3), n3))),
                         lambda o: (o.mscal(n3, n3, 0.5, C3, n3), o.gemm(111, 111, n3, n3, n3, 1.5, A3, n3, B3, n3,
                                                                         C3, n3))),
        "strsv_LNN_4096": ((nt * (nt + 1) // 2 + 2 * nt) * 4, "B",
                           lambda: so.exo_strsv_rm_stride_1(None, 122, 111, 131, nt, L.win1(P(xt), 1),
                                                            L.win2(P(T), nt)),
                           lambda o: o.trsv(122, 111, 131, nt, xt, 1, T, nt)),
    }
    o = Oracle(f32, kind="fast")
    kernels = {}
    for name, (work, unit, gpu, cpu) in cases.items():
        for _ in range(2):
            gpu()
        L.check()
        reps = 5 if work > 1e8 else 20
        t0 = time.perf_counter()
        for _ in range(reps):
            gpu()
        g_us = (time.perf_counter() - t0) / reps * 1e6
        cpu(o)
        c_reps = 1 if work > 1e8 else 3
        t0 = time.perf_counter()
        for _ in range(c_reps):
            cpu(o)
        c_us = (time.perf_counter() - t0) / c_reps * 1e6
        rate = (lambda us: round(work / (us * 1e-6) / 1e9, 2))
        kernels[name] = {"us": round(g_us, 1), f"G{unit}/s": rate(g_us), "cpu_us": round(c_us, 1),
                         f"cpu_G{unit}/s": rate(c_us), "speedup_vs_cpu": round(c_us / g_us, 2)}
    k = kernels["sgemm_NN_512"]
    return _line("drop-in host-pointer path: PCIe-inclusive rate per call", k["Gflop/s"], "GFLOP/s", a,
                 k["us"] / 1e3, "f32", "host buffers through the exo_* C-ABI: saxpy/sdot 2^26, sgemv N 8192^2, "
                 "sgemm NN 512^3 (mscal + gemm), strsv L-N-N 4096", None, kernels)


# ---------------------------------------------------------------- the other configurations
# The default run also measures every other configuration and the kernel
# families the bandwidth set leaves out, each with its own roofline and CPU
# baseline, under "configs" in the same JSON line.  Each section also runs
# alone with --workload <name> (for profiling).

def _u(torch, g, shape, dt):
    return torch.empty(shape, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)


def _roof(kernel, achieved, peak, unit, bound, traffic=None):
    return {"bound": bound, "kernel": kernel, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": traffic}


def _fl(v, p="s"):
    return ctypes.byref((ctypes.c_float if p == "s" else ctypes.c_double)(v))


def _timed_alt(torch, stream, fns, reps):
    """Average ms per launch of the callables in fns, taken in turn (each
    its own operand set), HIP events on `stream`."""
    for f in fns:
        f()
    return _timed(torch, stream, lambda it=iter(range(1 << 30)): fns[next(it) % len(fns)](), reps)


def _port(fn, work, div, unit, budget, sample, threads=1):
    sec, reps = cpu_rate(fn, budget)
    return {"value": round(work / sec / div, 4), "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{sample} ({reps} timed call(s), host-resident)"}


def _mkl_call(budget, threads, fn):
    mkl, name = _mkl()
    if mkl is None:
        return None, None
    mkl.MKL_Set_Num_Threads(ctypes.c_int(threads))
    sec, reps = cpu_rate(lambda: fn(mkl), budget)
    return sec, (name, reps)


def _vendor(budget, threads, fn, work, div, unit, sample):
    sec, info = _mkl_call(budget, threads, fn)
    if sec is None:
        return None
    return {"value": round(work / sec / div, 4), "unit": unit, "threads": threads,
            "sample": f"{sample}, {info[0]} ({info[1]} timed call(s))"}


def _ptr(x):
    return x.ctypes.data_as(ctypes.c_void_p)


def config1(a, torch, so, L):
    """Config 1: sgemm NN 512^3, A, B, C ~ U[-1,1] from seed 0, alpha 1.5,
    beta 0.5 applied as the wrapper does (mscal, then the product;
    test/level3/gemm/exo_gemm_wrapper.h:18-20), device-resident."""
    from oracle import Oracle
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    n = 512
    rng = np.random.default_rng(0)
    A, B, C = (rng.uniform(-1, 1, (n, n)).astype(np.float32) for _ in range(3))
    dA, dB, dC = (torch.from_numpy(t).cuda() for t in (A, B, C))
    wa, wb, wc = L.win2(dA.data_ptr(), n), L.win2(dB.data_ptr(), n), L.win2(dC.data_ptr(), n)

    def wrapper_sequence():
        r1 = so.exoblas_smscal_rm_stride_1_async(sh, n, n, _fl(0.5), wc)
        r2 = so.exoblas_sgemm_stride_1_async(sh, 111, 111, n, n, n, _fl(1.5), wa, wb, wa, wb, wc)
        return r1 or r2

    def fused():  # CBLAS-shaped extension: beta in the product's epilogue, one launch
        return so.exoblas_sgemm_ex_async(sh, 101, 111, 111, n, n, n, _fl(1.5), wa, wb, _fl(0.5), wc)
    for _ in range(5):
        wrapper_sequence()
        fused()
    ms_w = _timed_batch(torch, stream, wrapper_sequence, 10)
    ms = _timed_batch(torch, stream, fused, 10)
    flop = 2.0 * n ** 3
    tf, tf_w = flop / (ms * 1e-3) / 1e12, flop / (ms_w * 1e-3) / 1e12
    kernels = {"sgemm_NN_512_beta": {"us": round(ms * 1e3, 2), "TFLOP/s": round(tf, 2), "flop": flop,
                                     "launches": "one (exoblas_sgemm_ex_async, beta fused)",
                                     "timing": "device time per call, 20 calls back to back behind a spin"},
               "sgemm_NN_512_beta_wrapper": {"us": round(ms_w * 1e3, 2), "TFLOP/s": round(tf_w, 2), "flop": flop,
                                             "launches": "mscal + gemm, as the reference's wrapper calls them"}}
    out = {"workload": "sgemm NN 512^3, alpha 1.5, beta 0.5, seed 0 (beta fused, one launch)",
           "us": round(ms * 1e3, 2), "TFLOP/s": round(tf, 2),
           "roofline": _roof("sgemm_NN_512_beta", tf, FP32_PEAK_TF, "TFLOP/s", "mfma"), "kernels": kernels}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_leg("config1", a)
    return out


def _dd(torch, g, n, dt):
    A = _u(torch, g, (n, n), dt) / n
    A.diagonal().copy_(n + torch.empty(n, dtype=dt, device="cuda").uniform_(0, 1, generator=g))
    return A


def config3(a, torch, so, L):
    """Config 3: s/d trsv lower / no-trans / non-unit at n = 4096 (the
    diagonally dominant matrices, b ~ U[-1,1]; x reset before each timed
    solve, outside the timed region), s/d syrk lower-N N = 4096 K = 1024 and
    s/d symm left-lower M = N = 4096 (alpha 1.5)."""
    from oracle import Oracle
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    kernels = {}
    for p, dt, s_, peak in (("s", torch.float32, 4, FP32_PEAK_TF), ("d", torch.float64, 8, FP64_PEAK_TF)):
        n = 4096
        A = _dd(torch, g, n, dt)
        b = torch.empty(n, dtype=dt, device="cuda").uniform_(-1, 1, generator=g)
        x = b.clone()
        f = getattr(so, f"exoblas_{p}trsv_rm_stride_1_async")
        fn = (lambda f=f, A=A, x=x: f(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), L.win2(A.data_ptr(), n)))
        fn()
        ms = _timed(torch, stream, fn, 20, before=lambda x=x, b=b: x.copy_(b))
        nbytes = (n * (n + 1) // 2 + 2 * n) * s_
        gbs = nbytes / (ms * 1e-3) / 1e9
        kernels[f"{p}trsv_LNN_4096"] = {"us": round(ms * 1e3, 2), "GB/s": round(gbs, 1),
                                        "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": nbytes, "bound": "hbm"}
        del A, b, x
        N, K = 4096, 1024
        As, Cs = _u(torch, g, (N, K), dt), _u(torch, g, (N, N), dt)
        f = getattr(so, f"exoblas_{p}syrk_rm_stride_1_async")
        w = L.win2(As.data_ptr(), K)
        fn = (lambda f=f, w=w, Cs=Cs, p=p: f(sh, 122, 111, N, K, _fl(1.5, p), w, w, w, w, L.win2(Cs.data_ptr(), N)))
        fn()
        ms = _timed(torch, stream, fn, 10)
        flop = float(K) * N * (N + 1)
        tf = flop / (ms * 1e-3) / 1e12
        kernels[f"{p}syrk_LN_4096x1024"] = {"us": round(ms * 1e3, 2), "TFLOP/s": round(tf, 2),
                                            "frac": round(tf / peak, 4), "flop": flop, "bound": "mfma"}
        del As, Cs
        M = 4096
        Sa, Bm, Cm = (_u(torch, g, (M, M), dt) for _ in range(3))
        f = getattr(so, f"exoblas_{p}symm_rm_stride_1_async")
        fn = (lambda f=f, Sa=Sa, Bm=Bm, Cm=Cm, p=p: f(sh, 141, 122, M, M, _fl(1.5, p), L.win2(Sa.data_ptr(), M),
                                                      L.win2(Sa.data_ptr(), M), L.win2(Bm.data_ptr(), M),
                                                      L.win2(Cm.data_ptr(), M)))
        fn()
        ms = _timed(torch, stream, fn, 5)
        flop = 2.0 * M ** 3
        tf = flop / (ms * 1e-3) / 1e12
        kernels[f"{p}symm_LL_4096"] = {"us": round(ms * 1e3, 2), "TFLOP/s": round(tf, 2),
                                       "frac": round(tf / peak, 4), "flop": flop, "bound": "mfma"}
        del Sa, Bm, Cm
    torch.cuda.empty_cache()
    dom = max(kernels, key=lambda k: kernels[k]["us"])
    lowest = min(kernels, key=lambda k: kernels[k]["frac"])
    kd = kernels[dom]
    unit = "GB/s" if kd["bound"] == "hbm" else "TFLOP/s"
    peak_d = HBM_PEAK_GBS if kd["bound"] == "hbm" else (FP32_PEAK_TF if dom[0] == "s" else FP64_PEAK_TF)
    out = {"workload": "s/d trsv L-N-N 4096, s/d syrk L-N 4096x1024, s/d symm L-L 4096^2",
           "roofline": _roof(dom, kd[unit], peak_d, unit, kd["bound"]), "lowest_frac_kernel": lowest,
           "kernels": kernels}
    if not a.no_cpu:
        cpu = cpu_leg("config3", a)
        cb = dict(cpu.get(dom, {}))
        cb["kernels"] = {k: v for k, v in cpu.items() if k != "host"}
        cb["host"] = cpu.get("host")
        out["cpu_baseline"] = cb
    return out


def config4(a, torch, so, L):
    """Config 4: sgemm NN n^3 (n = --n-gemm, 8192) plain FP32 MFMA against the
    3-, 4- and 6-product split-BF16 forms, on U[-1,1] and on rows scaled by
    2^U[-8,8], with both error metrics against float64 (max |C - C64| /
    (|A||B|) and max |C - C64| / |C64|); dgemm NN n^3; ssyrk L-N n x n/4 plain
    against the splits; strsv L-N-N n plain against the mixed solve (BF16 /
    FP32 off-diagonal updates, FP64 diagonal blocks, one refinement step)."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    n = a.n_gemm
    kernels = {}
    one = _fl(1.0)
    forms = ("plain", 3, 4, 6)

    def peak_of(form):
        return FP32_PEAK_TF if form == "plain" else BF16_PEAK_TF / form

    def errs(C, truth, mag, mask=None):
        err = (C.double() - truth).abs()
        nw, rel = err / mag, err / truth.abs().clamp_min(1e-300)
        if mask is not None:
            nw, rel = nw[mask], rel[mask]
        return nw.max().item(), rel.max().item()
    for scaled in (False, True):
        A = _u(torch, g, (n, n), torch.float32)
        B = _u(torch, g, (n, n), torch.float32)
        if scaled:
            A *= torch.exp2(torch.empty((n, 1), device="cuda").uniform_(-8, 8, generator=g))
            B *= torch.exp2(torch.empty((n, 1), device="cuda").uniform_(-8, 8, generator=g))
        truth = A.double() @ B.double()
        mag = A.double().abs() @ B.double().abs()
        C = torch.zeros((n, n), device="cuda")
        wa, wb, wc = L.win2(A.data_ptr(), n), L.win2(B.data_ptr(), n), L.win2(C.data_ptr(), n)
        for form in forms:
            if form == "plain":
                fn = (lambda: so.exoblas_sgemm_stride_1_async(sh, 111, 111, n, n, n, one, wa, wb, wa, wb, wc))
            else:
                fn = (lambda form=form: so.exoblas_sgemm_split_async(sh, 111, 111, n, n, n, one, wa, wb, wa, wb,
                                                                      wc, form))
            C.zero_()
            assert fn() == 0, L.last_error()
            torch.cuda.synchronize()
            nw, rel = errs(C, truth, mag)
            key = f"sgemm_{form if form == 'plain' else 'split' + str(form)}"
            e = kernels.setdefault(key, {"terms": None if form == "plain" else form})
            e[f"max_err_normwise_{'scaled' if scaled else 'uniform'}"] = nw
            e[f"max_rel_err_{'scaled' if scaled else 'uniform'}"] = rel
            if not scaled:
                ms = _timed(torch, stream, fn, 3)
                tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
                e.update({"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "frac": round(tf / peak_of(form), 4),
                          "peak_equiv": round(peak_of(form), 1)})
        del truth, mag, A, B, C
        torch.cuda.empty_cache()
    # dgemm NN n^3
    A, B, C = (_u(torch, g, (n, n), torch.float64) for _ in range(3))
    w = [L.win2(t.data_ptr(), n) for t in (A, B, C)]
    fn = (lambda: so.exoblas_dgemm_stride_1_async(sh, 111, 111, n, n, n, _fl(1.5, "d"), w[0], w[1], w[0], w[1], w[2]))
    fn()
    ms = _timed(torch, stream, fn, 2)
    tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
    kernels["dgemm_plain"] = {"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "frac": round(tf / FP64_PEAK_TF, 4)}
    del A, B, C
    torch.cuda.empty_cache()
    # ssyrk lower-N, N = n, K = n / 4
    N, K = n, n // 4
    tri = torch.ones((N, N), dtype=torch.bool, device="cuda").tril_()
    for scaled in (False, True):
        A = _u(torch, g, (N, K), torch.float32)
        if scaled:
            A *= torch.exp2(torch.empty((N, 1), device="cuda").uniform_(-8, 8, generator=g))
        A64 = A.double()
        truth = A64 @ A64.T
        mag = A64.abs() @ A64.abs().T
        del A64
        C = torch.zeros((N, N), device="cuda")
        wa, wc = L.win2(A.data_ptr(), K), L.win2(C.data_ptr(), N)
        for form in forms:
            if form == "plain":
                fn = (lambda: so.exoblas_ssyrk_rm_stride_1_async(sh, 122, 111, N, K, one, wa, wa, wa, wa, wc))
            else:
                fn = (lambda form=form: so.exoblas_ssyrk_split_async(sh, 122, 111, N, K, one, wa, wa, wc, form))
            C.zero_()
            assert fn() == 0, L.last_error()
            torch.cuda.synchronize()
            nw, rel = errs(C, truth, mag, tri)
            key = f"ssyrk_{form if form == 'plain' else 'split' + str(form)}"
            e = kernels.setdefault(key, {"terms": None if form == "plain" else form})
            e[f"max_err_normwise_{'scaled' if scaled else 'uniform'}"] = nw
            e[f"max_rel_err_{'scaled' if scaled else 'uniform'}"] = rel
            if not scaled:
                ms = _timed(torch, stream, fn, 3)
                flop = float(K) * N * (N + 1)
                tf = flop / (ms * 1e-3) / 1e12
                e.update({"ms": round(ms, 3), "TFLOP/s": round(tf, 2), "frac": round(tf / peak_of(form), 4)})
        del A, C, truth, mag
        torch.cuda.empty_cache()
    del tri
    # strsv n: plain vs mixed (+ one refinement step)
    At = _dd(torch, g, n, torch.float32)
    b = torch.empty(n, device="cuda").uniform_(-1, 1, generator=g)
    x64 = torch.linalg.solve_triangular(At.double().tril(), b.double()[:, None], upper=False)[:, 0]
    x = b.clone()
    wA = L.win2(At.data_ptr(), n)
    packed = torch.empty(n * n, dtype=torch.bfloat16, device="cuda")  # the BF16 copy, made once
    assert so.exoblas_strsv_mixed_pack_async(sh, n, wA, packed.data_ptr()) == 0, L.last_error()
    tforms = {"plain": lambda: so.exoblas_strsv_rm_stride_1_async(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA),
              "mixed_bf16": lambda: so.exoblas_strsv_mixed_async(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA,
                                                                 16, 1),
              "mixed_bf16_packed": lambda: so.exoblas_strsv_mixed_packed_async(
                  sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA, packed.data_ptr(), 1),
              "mixed_fp32": lambda: so.exoblas_strsv_mixed_async(sh, 122, 111, 131, n, L.win1(x.data_ptr(), 1), wA,
                                                                 32, 1)}
    nbytes = (n * (n + 1) // 2 + 2 * n) * 4
    for name, fn in tforms.items():
        x.copy_(b)
        assert fn() == 0, L.last_error()
        torch.cuda.synchronize()
        err = (x.double() - x64).abs()
        ms = _timed(torch, stream, fn, 10, before=lambda: x.copy_(b))
        kernels[f"strsv_{name}"] = {"us": round(ms * 1e3, 2), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                                    "max_err": err.max().item(),
                                    "max_rel_err": (err / x64.abs().clamp_min(1e-300)).max().item()}
    del At, b, x, x64, packed
    torch.cuda.empty_cache()
    plain = kernels["sgemm_plain"]

    def le_plain(key, base):
        return {f"{m}_{st}": kernels[key][f"{m}_{st}"] <= kernels[base][f"{m}_{st}"]
                for m in ("max_err_normwise", "max_rel_err") for st in ("uniform", "scaled")}
    summary = {f: {"speedup": round(plain["ms"] / kernels[f"sgemm_split{f}"]["ms"], 3),
                   "error_le_plain": le_plain(f"sgemm_split{f}", "sgemm_plain"),
                   "ssyrk_speedup": round(kernels["ssyrk_plain"]["ms"] / kernels[f"ssyrk_split{f}"]["ms"], 3),
                   "ssyrk_error_le_plain": le_plain(f"ssyrk_split{f}", "ssyrk_plain")} for f in (3, 4, 6)}
    # per error metric, the fastest form whose sgemm and ssyrk errors are at
    # most plain FP32's on both sets (max relative error is the stated
    # precision metric; the normwise error |C - C64| / (|A||B|) is the other)
    claims = {}
    for m in ("max_rel_err", "max_err_normwise"):
        ok = [f for f in (3, 4, 6) if all(summary[f]["error_le_plain"][f"{m}_{st}"] and
                                          summary[f]["ssyrk_error_le_plain"][f"{m}_{st}"]
                                          for st in ("uniform", "scaled"))]
        best = max(ok, key=lambda f: summary[f]["speedup"]) if ok else None
        claims[m] = {"form": f"{best} products" if best else None,
                     "sgemm_speedup": summary[best]["speedup"] if best else None,
                     "ssyrk_speedup": summary[best]["ssyrk_speedup"] if best else None}
    out = {"workload": f"sgemm NN {n}^3 plain vs split-BF16 (3/4/6 products), uniform and row-scaled sets; dgemm NN "
                       f"{n}^3; ssyrk L-N {N}x{K} plain vs split; strsv L-N-N {n} plain vs mixed + 1 refinement",
           "roofline": _roof("sgemm_plain", plain["TFLOP/s"], FP32_PEAK_TF, "TFLOP/s", "mfma"),
           "split_forms": summary, "precision_claims": claims, "kernels": kernels}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_leg("config4", a)
    return out


def config5(a, torch, so, L):
    """Config 5: the seeded mixed request stream (60 % sgemm NN, 25 % strsv,
    15 % sgemv N), batched persistent kernels against one launch per call,
    64 GB admission budget.  Every request reads its own window of 512 MB
    operand pools (gemm A and B, gemv A) or one of 24 2048^2 triangular
    matrices, and writes its own output.  Capacity from all-at-once arrivals;
    latency under load from Poisson arrivals at rates from below to well
    above the per-call capacity (the same request sequence, arrival gaps
    scaled)."""
    from exoblas_amd import serving
    s = serving.make_stream(a.requests, rate=a.rate, seed=0)
    w = serving.Workload(s, pool_bytes=512 << 20)
    res = {}

    def one(mode, arrival):
        w.reset()
        torch.cuda.synchronize()
        st, _ = serving.run(w, mode=mode, budget=64 << 30, streams=4, max_batch=a.max_batch, arrival_us=arrival)
        return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}
    for mode, name in ((1, "per_call"), (0, "batched")):
        res[f"burst_{name}"] = one(mode, np.zeros(a.requests))
    sweep = {}
    for rate in (20000.0, 50000.0, 100000.0, 200000.0, 400000.0, 600000.0):
        arr = s.arrival_us * (a.rate / rate)
        sweep[f"{rate:.0f}"] = {name: {k: v for k, v in one(mode, arr).items()
                                       if k in ("req_per_s", "p50_us", "p99_us", "mean_us", "completed", "batches",
                                                "peak_bytes_in_flight")}
                                for mode, name in ((1, "per_call"), (0, "batched"))}
    at_rate = sweep.get(f"{a.rate:.0f}") or {n: one(m, s.arrival_us) for m, n in ((1, "per_call"), (0, "batched"))}
    cap = res["burst_batched"]["req_per_s"]
    out = {"workload": f"{a.requests} requests, 60% sgemm / 25% strsv / 15% sgemv, per-request operand windows of "
                       f"512 MB pools; all-at-once arrivals and Poisson arrivals at 20k-600k/s",
           "req_per_s": round(cap, 1),
           "speedup_vs_per_call": round(cap / max(res["burst_per_call"]["req_per_s"], 1e-9), 3),
           "p50_us": at_rate["batched"]["p50_us"], "p99_us": at_rate["batched"]["p99_us"],
           "p50_us_per_call": at_rate["per_call"]["p50_us"], "p99_us_per_call": at_rate["per_call"]["p99_us"],
           "roofline": {"bound": "mixed", "kernel": "batch_kernel", "achieved": None, "peak": None, "unit": None,
                        "frac": None, "traffic": None},
           "modes": res, "poisson_sweep": sweep}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_leg("config5", a)
    del w
    torch.cuda.empty_cache()
    return out


def serving_cpu_baseline(s, budget):
    """The same request stream, front to back, through the oracle's loop
    nests on one host thread (gemm as mscal + gemm): requests per second over
    as many requests as the budget allows."""
    from exoblas_amd import native as N
    from oracle import Oracle
    o = Oracle(np.float32, kind="fast")
    rng = np.random.default_rng(1)
    pool = rng.uniform(-1, 1, 1 << 26).astype(np.float32)  # operand data, drawn before timing
    done, busy = 0, 0.0
    for i in range(len(s.kind)):
        m, n, k = int(s.m[i]), int(s.n[i]), int(s.k[i])
        if s.kind[i] == N.OP_GEMM:
            A, B = pool[:m * k].reshape(m, k), pool[:k * n].reshape(k, n)
            C = pool[-m * n:].reshape(m, n).copy()
            t0 = time.perf_counter()
            o.mscal(m, n, 0.5, C, n)
            o.gemm(111, 111, m, n, k, 1.5, A, k, B, n, C, n)
        elif s.kind[i] == N.OP_TRSV:
            T = pool[:m * m].reshape(m, m) / m
            T[np.arange(m), np.arange(m)] = m + 0.5
            x = pool[:m].copy()
            t0 = time.perf_counter()
            o.trsv(122, 111, 131, m, x, 1, T, m)
        else:
            A = pool[:m * n].reshape(m, n)
            y = pool[-m:].copy()
            t0 = time.perf_counter()
            o.gemv(111, m, n, 1.5, 0.5, A, n, pool[:n], 1, y, 1)
        busy += time.perf_counter() - t0
        done += 1
        if busy > budget:
            break
    el = busy
    return {"value": round(done / el, 2), "unit": "req/s", "cores": 1, "kind": "port",
            "sample": f"the first {done} requests of the same stream in order, oracle loop nests, 1 thread"}


def families(a, torch, so, L):
    """The hot-path kernel families the bandwidth set leaves out, each at a
    bandwidth-sized shape with two operand sets taken in turn (so no call
    reads what the previous one touched): s/d asum, dsdot, sdsdot, s/d rotm
    (flag -1), s/d scal at n = 2^26; s/d ger 8192^2; ssymv U and L, ssyr,
    ssyr2, strmv L-N-N at n = 8192; s/d mscal and strmscal 8192^2.
    GB/s over the compulsory bytes (s = element size): asum ns, dsdot and
    sdsdot 8n, rotm 4ns, scal 2ns, ger 2MNs + (M+N)s, symv n(n+1)/2 s + 3ns,
    syr n(n+1) s + ns, syr2 n(n+1) s + 2ns, trmv n(n+1)/2 s + 2ns, mscal
    2MNs, trmscal n(n+1) s."""
    from oracle import Oracle
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    n1, n2 = a.n1, a.n2
    kernels = {}
    w1 = lambda t: L.win1(t.data_ptr(), 1)
    w2 = lambda t, ld: L.win2(t.data_ptr(), ld)
    res = torch.zeros(2, dtype=torch.float64, device="cuda")
    rp = ctypes.c_void_p(res.data_ptr())

    def run(name, make, nbytes, reps=10):
        fns = [make(k) for k in range(2)]
        ms = _timed_alt(torch, stream, fns, reps)
        gbs = nbytes / (ms * 1e-3) / 1e9
        kernels[name] = {"us": round(ms * 1e3, 2), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "bytes": nbytes}
    for p, dt, s_ in (("s", torch.float32, 4), ("d", torch.float64, 8)):
        X = [_u(torch, g, n1, dt) for _ in range(2)]
        Y = [_u(torch, g, n1, dt) for _ in range(2)]
        run(f"{p}asum", lambda k: (lambda: getattr(so, f"exoblas_{p}asum_stride_1_async")(sh, n1, w1(X[k]), rp)),
            n1 * s_)
        H = (ctypes.c_float * 4)(0.6, 0.8, -0.8, 0.6) if p == "s" else (ctypes.c_double * 4)(0.6, 0.8, -0.8, 0.6)
        run(f"{p}rotm_neg_one", lambda k: (lambda: getattr(so, f"exoblas_{p}rotm_flag_neg_one_stride_1_async")(
            sh, n1, w1(X[k]), w1(Y[k]), H)), 4 * n1 * s_)
        run(f"{p}scal", lambda k: (lambda: getattr(so, f"exoblas_{p}scal_stride_1_async")(
            sh, n1, _fl(0.999, p), w1(X[k]))), 2 * n1 * s_)
        if p == "s":
            run("dsdot", lambda k: (lambda: so.exoblas_dsdot_stride_1_async(sh, n1, w1(X[k]), w1(Y[k]), rp)), 8 * n1)
            run("sdsdot", lambda k: (lambda: so.exoblas_sdsdot_stride_1_async(sh, n1, _fl(0.25), w1(X[k]), w1(Y[k]),
                                                                            rp)), 8 * n1)
        del X, Y
        Am = [_u(torch, g, (n2, n2), dt) for _ in range(2)]
        xv = [_u(torch, g, n2, dt) for _ in range(2)]
        yv = [_u(torch, g, n2, dt) for _ in range(2)]
        run(f"{p}ger", lambda k: (lambda: getattr(so, f"exoblas_{p}ger_rm_stride_1_async")(
            sh, n2, n2, _fl(1e-3, p), w1(xv[k]), w1(yv[k]), w2(Am[k], n2))), 2 * n2 * n2 * s_ + 2 * n2 * s_)
        run(f"{p}mscal", lambda k: (lambda: getattr(so, f"exoblas_{p}mscal_rm_stride_1_async")(
            sh, n2, n2, _fl(0.999, p), w2(Am[k], n2))), 2 * n2 * n2 * s_)
        if p == "s":
            tri = n2 * (n2 + 1) // 2
            for u, uname in ((121, "U"), (122, "L")):
                run(f"ssymv_{uname}", lambda k, u=u: (lambda: so.exoblas_ssymv_rm_stride_1_async(
                    sh, u, n2, _fl(0.5), w2(Am[k], n2), w1(xv[k]), _fl(1.5), w1(yv[k]))), tri * 4 + 3 * n2 * 4)
            run("ssyr_L", lambda k: (lambda: so.exoblas_ssyr_rm_stride_1_async(
                sh, 122, n2, _fl(1e-3), w1(xv[k]), w1(xv[k]), w2(Am[k], n2))), 2 * tri * 4 + n2 * 4)
            run("ssyr2_L", lambda k: (lambda: so.exoblas_ssyr2_rm_stride_1_async(
                sh, 122, n2, _fl(1e-3), w1(xv[k]), w1(xv[k]), w1(yv[k]), w1(yv[k]), w2(Am[k], n2))),
                2 * tri * 4 + 2 * n2 * 4)
            run("strmv_LNN", lambda k: (lambda: so.exoblas_strmv_rm_stride_1_async(
                sh, 122, 111, 131, n2, w1(xv[k]), w2(Am[k], n2))), tri * 4 + 2 * n2 * 4)
            run("strmscal_L", lambda k: (lambda: so.exoblas_strmscal_rm_stride_1_async(
                sh, 122, n2, _fl(0.999), w2(Am[k], n2))), 2 * tri * 4)
        del Am, xv, yv
        torch.cuda.empty_cache()
    lowest = min(kernels, key=lambda k: kernels[k]["frac"])
    dom = max(kernels, key=lambda k: kernels[k]["us"])
    out = {"workload": "s/d asum, dsdot, sdsdot, s/d rotm, s/d scal n=2^26; s/d ger, s/d mscal 8192^2; ssymv U/L, "
                       "ssyr, ssyr2, strmv, strmscal n=8192",
           "roofline": _roof(dom, kernels[dom]["GB/s"], HBM_PEAK_GBS, "GB/s", "hbm"), "lowest_frac_kernel": lowest,
           "kernels": kernels}
    if not a.no_cpu:
        cpu = cpu_leg("families", a)
        cb = dict(cpu.get(dom, cpu.get("sger", {})))
        cb["kernels"] = cpu
        out["cpu_baseline"] = cb
    return out



def strided(a, torch, so, L):
    """Strided and misaligned windows through the element-by-element kernels:
    s/d axpy, sdot, scopy at incx = 2, incy = -3 and n = 2^26, saxpy on
    unit-stride windows one element off 16-byte alignment, and sgemv N / T
    8192^2 with incx = 2, incy = -3 (two operand sets taken in turn).  Two
    byte counts per kernel: `bytes`, the elements moved (as for unit stride:
    axpy 3ns, dot 2ns, copy 2ns, gemv (n^2 + 3n)s), and `touched`, the
    memory lines the elements lie in (each window's span |inc| n s, counted
    twice when written), which is what HBM has to move; `frac` is GB/s over
    touched bytes against HBM peak."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    n1, n2 = a.n1, a.n2
    kernels = {}
    res = torch.zeros(2, dtype=torch.float64, device="cuda")
    rp = ctypes.c_void_p(res.data_ptr())

    def span(n, inc, s_):
        return n * min(abs(inc) * s_, 128)

    def win(t, n, inc, s_):  # element 0 of an n-vector at increment inc in buffer t
        return L.win1(t.data_ptr() + ((n - 1) * abs(inc) * s_ if inc < 0 else 0), inc)

    def run(name, make, nbytes, touched, reps=10):
        fns = [make(k) for k in range(2)]
        ms = _timed_alt(torch, stream, fns, reps)
        gbs, gbt = nbytes / (ms * 1e-3) / 1e9, touched / (ms * 1e-3) / 1e9
        kernels[name] = {"us": round(ms * 1e3, 2), "GB/s": round(gbs, 1), "GB/s_touched": round(gbt, 1),
                         "frac": round(gbt / HBM_PEAK_GBS, 4), "bytes": nbytes, "touched": touched}
    ix, iy = 2, -3
    for p, dt, s_ in (("s", torch.float32, 4), ("d", torch.float64, 8)):
        X = [_u(torch, g, 1 + (n1 - 1) * abs(ix), dt) for _ in range(2)]
        Y = [_u(torch, g, 1 + (n1 - 1) * abs(iy), dt) for _ in range(2)]
        t_xy = span(n1, ix, s_) + 2 * span(n1, iy, s_)
        run(f"{p}axpy_inc2_-3", lambda k: (lambda: getattr(so, f"exoblas_{p}axpy_stride_any_async")(
            sh, n1, _fl(0.5, p), win(X[k], n1, ix, s_), win(Y[k], n1, iy, s_))), 3 * n1 * s_, t_xy)
        if p == "s":
            run("sdot_inc2_-3", lambda k: (lambda: so.exoblas_sdot_stride_any_async(
                sh, n1, win(X[k], n1, ix, 4), win(Y[k], n1, iy, 4), rp)), 2 * n1 * 4,
                span(n1, ix, 4) + span(n1, iy, 4))
            run("scopy_inc2_-3", lambda k: (lambda: so.exoblas_scopy_stride_any_async(
                sh, n1, win(X[k], n1, ix, 4), win(Y[k], n1, iy, 4))), 2 * n1 * 4, span(n1, ix, 4) + 2 * span(n1, iy, 4))
        del X, Y
        torch.cuda.empty_cache()
    Xm = [_u(torch, g, n1 + 1, torch.float32) for _ in range(2)]
    Ym = [_u(torch, g, n1 + 1, torch.float32) for _ in range(2)]
    run("saxpy_misaligned", lambda k: (lambda: so.exoblas_saxpy_stride_any_async(
        sh, n1, _fl(0.5), L.win1(Xm[k].data_ptr() + 4, 1), L.win1(Ym[k].data_ptr() + 4, 1))), 12 * n1, 12 * n1)
    del Xm, Ym
    Am = [_u(torch, g, (n2, n2), torch.float32) for _ in range(2)]
    xv = [_u(torch, g, 1 + (n2 - 1) * abs(ix), torch.float32) for _ in range(2)]
    yv = [_u(torch, g, 1 + (n2 - 1) * abs(iy), torch.float32) for _ in range(2)]
    gb = (n2 * n2 + 3 * n2) * 4
    gt = n2 * n2 * 4 + span(n2, ix, 4) + 2 * span(n2, iy, 4)
    for tr, tname in ((111, "N"), (112, "T")):
        run(f"sgemv_{tname}_inc2_-3", lambda k, tr=tr: (lambda: so.exoblas_sgemv_rm_stride_any_async(
            sh, tr, n2, n2, _fl(1.5), _fl(0.5), L.win2(Am[k].data_ptr(), n2), L.win2(Am[k].data_ptr(), n2),
            win(xv[k], n2, ix, 4), win(yv[k], n2, iy, 4))), gb, gt)
    del Am, xv, yv
    torch.cuda.empty_cache()
    lowest = min(kernels, key=lambda k: kernels[k]["frac"])
    dom = max(kernels, key=lambda k: kernels[k]["us"])
    out = {"workload": "s/d axpy, sdot, scopy incx=2 incy=-3 n=2^26; saxpy misaligned n=2^26; sgemv N/T 8192^2 "
                       "incx=2 incy=-3",
           "roofline": _roof(dom, kernels[dom]["GB/s_touched"], HBM_PEAK_GBS, "GB/s", "hbm"),
           "lowest_frac_kernel": lowest, "kernels": kernels}
    if not a.no_cpu:
        cpu = cpu_leg("strided", a)
        cb = dict(cpu.get("saxpy_inc2_-3", {}))
        cb["kernels"] = cpu
        out["cpu_baseline"] = cb
    return out


# ---------------------------------------------------------------- CPU legs
# Every CPU baseline runs in a child process that never imports torch: with
# torch's OpenMP runtime in the process the vendor CBLAS's own threads
# contend with it (MKL sgemm 2048^3 measured 963 GFLOP/s on 8 threads alone,
# 108 GFLOP/s with torch imported, below its single-thread 133).

def cpu_leg(name, a):
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", name, "--cpu-budget", str(a.cpu_budget),
           "--n1", str(a.n1), "--n2", str(a.n2), "--n-gemm", str(a.n_gemm), "--requests", str(a.requests),
           "--rate", str(a.rate)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return {"error": f"cpu leg {name} failed: {r.stderr[-400:]}"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_bandwidth(a):
    hc = host_cpus()
    cb = cpu_baseline(a.n1, a.n2, a.cpu_budget)
    cb["host"] = hc
    set_port_threads(hc["threads"])
    cb["all_cores"] = cpu_baseline(a.n1, a.n2, a.cpu_budget / 3, kind="omp", threads=hc["threads"])
    vend = [vendor_cblas(a.n1, a.n2, a.cpu_budget / 4, t) for t in sorted({1, hc["threads"]})]
    if vend[0] is not None:
        cb["vendor_cblas"] = vend
    return cb


def cpu_config1(a):
    from oracle import Oracle
    n = 512
    flop = 2.0 * n ** 3
    rng = np.random.default_rng(0)
    A, B, C = (rng.uniform(-1, 1, (n, n)).astype(np.float32) for _ in range(3))
    hc = host_cpus()
    T = hc["threads"]
    o = Oracle(np.float32, kind="fast")
    Cc = C.copy()
    seq = (lambda o=o: (o.mscal(n, n, 0.5, Cc, n), o.gemm(111, 111, n, n, n, 1.5, A, n, B, n, Cc, n)))
    cb = _port(seq, flop, 1e9, "GFLOP/s", a.cpu_budget / 5, "the full workload: the oracle's mscal + gemm "
               "loop nests, avx512 -O3 -ffast-math, 1 thread")
    set_port_threads(T)
    oo = Oracle(np.float32, kind="omp")
    seq2 = (lambda o=oo: (o.mscal(n, n, 0.5, Cc, n), o.gemm(111, 111, n, n, n, 1.5, A, n, B, n, Cc, n)))
    cb["all_cores"] = _port(seq2, flop, 1e9, "GFLOP/s", a.cpu_budget / 10, f"the same, OpenMP over rows",
                            threads=T)

    def mk(mkl):
        mkl.cblas_sgemm(101, 111, 111, n, n, n, ctypes.c_float(1.5), _ptr(A), n, _ptr(B), n,
                        ctypes.c_float(0.5), _ptr(Cc), n)
    cb["vendor_cblas"] = [v for v in (_vendor(a.cpu_budget / 10, t, mk, flop, 1e9, "GFLOP/s",
                                              "cblas_sgemm 512^3 with beta") for t in sorted({1, T})) if v]
    cb["host"] = hc
    return cb


def cpu_config3(a):
    from oracle import Oracle
    hc = host_cpus()
    T = hc["threads"]
    rng = np.random.default_rng(0)
    cpu = {}
    for p, dt in (("s", np.float32), ("d", np.float64)):
        o = Oracle(dt, kind="fast")
        n = 4096
        At = (rng.uniform(-1, 1, (n, n)) / n).astype(dt)
        At[np.arange(n), np.arange(n)] = (n + rng.uniform(0, 1, n)).astype(dt)
        bt = rng.uniform(-1, 1, n).astype(dt)
        nb = (n * (n + 1) // 2 + 2 * n) * np.dtype(dt).itemsize
        cpu[f"{p}trsv_LNN_4096"] = _port(lambda o=o, At=At, bt=bt: o.trsv(122, 111, 131, n, bt.copy(), 1, At, n),
                                         nb, 1e9, "GB/s", a.cpu_budget / 20, "the full solve, 1 thread")
        N, K = 1024, 256
        As = rng.uniform(-1, 1, (N, K)).astype(dt)
        Cs = rng.uniform(-1, 1, (N, N)).astype(dt)
        cpu[f"{p}syrk_LN_4096x1024"] = _port(lambda o=o, As=As, Cs=Cs: o.syrk(122, 111, N, K, 1.5, As, K, Cs, N),
                                             float(K) * N * (N + 1), 1e9, "GFLOP/s", a.cpu_budget / 20,
                                             "rate on syrk L-N 1024x256, 1 thread")
        M = 512
        Sa, Bm, Cm = (rng.uniform(-1, 1, (M, M)).astype(dt) for _ in range(3))
        cpu[f"{p}symm_LL_4096"] = _port(lambda o=o, Sa=Sa, Bm=Bm, Cm=Cm: o.symm(141, 122, M, M, 1.5, Sa, M, Bm, M,
                                                                             Cm, M),
                                        2.0 * M ** 3, 1e9, "GFLOP/s", a.cpu_budget / 20,
                                        "rate on symm L-L 512^2, 1 thread")
        # the vendor CBLAS at the full sizes
        N, K, M = 4096, 1024, 4096
        Af = rng.uniform(-1, 1, (N, K)).astype(dt)
        Cf = rng.uniform(-1, 1, (N, N)).astype(dt)
        Sf, Bf = (rng.uniform(-1, 1, (M, M)).astype(dt) for _ in range(2))
        Tc = ctypes.c_float if p == "s" else ctypes.c_double
        vend = {}
        for t in sorted({1, T}):
            vs = _vendor(a.cpu_budget / 30, t, lambda m, Af=Af, Cf=Cf: getattr(m, f"cblas_{p}syrk")(
                101, 122, 111, N, K, Tc(1.5), _ptr(Af), K, Tc(1.0), _ptr(Cf), N), float(K) * N * (N + 1), 1e9,
                "GFLOP/s", f"cblas_{p}syrk L-N 4096x1024")
            vm = _vendor(a.cpu_budget / 30, t, lambda m, Sf=Sf, Bf=Bf, Cf=Cf: getattr(m, f"cblas_{p}symm")(
                101, 141, 122, M, M, Tc(1.5), _ptr(Sf), M, _ptr(Bf), M, Tc(1.0), _ptr(Cf), M), 2.0 * M ** 3, 1e9,
                "GFLOP/s", f"cblas_{p}symm L-L 4096^2")
            vt = _vendor(a.cpu_budget / 60, t, lambda m, At=At, bt=bt: getattr(m, f"cblas_{p}trsv")(
                101, 122, 111, 131, n, _ptr(At), n, _ptr(bt.copy()), 1), (n * (n + 1) // 2 + 2 * n) *
                np.dtype(dt).itemsize, 1e9, "GB/s", f"cblas_{p}trsv L-N-N 4096")
            vend[t] = {"syrk": vs, "symm": vm, "trsv": vt}
        cpu[f"{p}_vendor_cblas"] = vend
    cpu["host"] = hc
    return cpu


def cpu_config4(a):
    hc = host_cpus()
    T = hc["threads"]
    rng = np.random.default_rng(0)
    m = 2048
    Am, Bm, Cm = (rng.uniform(-1, 1, (m, m)).astype(np.float32) for _ in range(3))
    vend = [v for v in (_vendor(a.cpu_budget / 10, t, lambda mk: mk.cblas_sgemm(
        101, 111, 111, m, m, m, ctypes.c_float(1.0), _ptr(Am), m, _ptr(Bm), m, ctypes.c_float(0.0), _ptr(Cm), m),
        2.0 * m ** 3, 1e9, "GFLOP/s", "rate on cblas_sgemm NN 2048^3") for t in sorted({1, T})) if v]
    from oracle import Oracle
    o = Oracle(np.float32, kind="fast")
    p5 = 512
    A5, B5, C5 = (rng.uniform(-1, 1, (p5, p5)).astype(np.float32) for _ in range(3))
    cb = _port(lambda: o.gemm(111, 111, p5, p5, p5, 1.0, A5, p5, B5, p5, C5, p5), 2.0 * p5 ** 3, 1e9, "GFLOP/s",
               a.cpu_budget / 10, "rate of the oracle's gemm loop nest on 512^3, 1 thread")
    cb["vendor_cblas"] = vend
    cb["host"] = hc
    return cb


def cpu_config5(a):
    from exoblas_amd import serving
    return serving_cpu_baseline(serving.make_stream(a.requests, rate=a.rate, seed=0), a.cpu_budget / 5)


def cpu_families(a):
    from oracle import Oracle
    rng = np.random.default_rng(0)
    o = Oracle(np.float32, kind="fast")
    m1, m2 = a.n1, a.n2
    x, y = rng.uniform(-1, 1, m1).astype(np.float32), rng.uniform(-1, 1, m1).astype(np.float32)
    A = rng.uniform(-1, 1, (m2, m2)).astype(np.float32)
    xv, yv = rng.uniform(-1, 1, m2).astype(np.float32), rng.uniform(-1, 1, m2).astype(np.float32)
    tri = m2 * (m2 + 1) // 2
    b = a.cpu_budget / 40
    cpu = {
        "sasum": _port(lambda: o.asum(m1, x, 1), m1 * 4, 1e9, "GB/s", b, "sasum n=2^26, 1 thread"),
        "dsdot": _port(lambda: o.dsdot(m1, x, 1, y, 1), 8 * m1, 1e9, "GB/s", b, "dsdot n=2^26, 1 thread"),
        "srotm_neg_one": _port(lambda: o.rotm(-1, m1, x, 1, y, 1, np.array([0.6, 0.8, -0.8, 0.6], np.float32)),
                               16 * m1, 1e9, "GB/s", b, "srotm n=2^26, 1 thread"),
        "sscal": _port(lambda: o.scal(m1, 0.999, x, 1), 8 * m1, 1e9, "GB/s", b, "sscal n=2^26, 1 thread"),
        "sger": _port(lambda: o.ger(m2, m2, 1e-3, xv, 1, yv, 1, A, m2), 8 * m2 * m2, 1e9, "GB/s", b,
                      "sger 8192^2, 1 thread"),
        "ssymv_L": _port(lambda: o.symv(122, m2, 0.5, A, m2, xv, 1, 1.5, yv.copy(), 1), 4 * tri, 1e9, "GB/s", b,
                         "ssymv L 8192, 1 thread"),
        "ssyr_L": _port(lambda: o.syr(122, m2, 1e-3, xv, 1, A, m2), 8 * tri, 1e9, "GB/s", b,
                        "ssyr L 8192, 1 thread"),
        "ssyr2_L": _port(lambda: o.syr2(122, m2, 1e-3, xv, 1, yv, 1, A, m2), 8 * tri, 1e9, "GB/s", b,
                         "ssyr2 L 8192, 1 thread"),
        "strmv_LNN": _port(lambda: o.trmv(122, 111, 131, m2, xv.copy(), 1, A, m2), 4 * tri, 1e9, "GB/s", b,
                           "strmv L-N-N 8192, 1 thread"),
        "smscal": _port(lambda: o.mscal(m2, m2, 0.999, A, m2), 8 * m2 * m2, 1e9, "GB/s", b,
                        "smscal 8192^2, 1 thread"),
    }
    return cpu


def cpu_strided(a):
    from oracle import Oracle
    rng = np.random.default_rng(0)
    o = Oracle(np.float32, kind="fast")
    m1, m2 = a.n1, a.n2
    x = rng.uniform(-1, 1, 1 + (m1 - 1) * 2).astype(np.float32)
    y = rng.uniform(-1, 1, 1 + (m1 - 1) * 3).astype(np.float32)
    A = rng.uniform(-1, 1, (m2, m2)).astype(np.float32)
    xv = rng.uniform(-1, 1, 1 + (m2 - 1) * 2).astype(np.float32)
    yv = rng.uniform(-1, 1, 1 + (m2 - 1) * 3).astype(np.float32)
    b = a.cpu_budget / 12
    return {
        "saxpy_inc2_-3": _port(lambda: o.axpy(m1, 0.5, x, 2, y, -3), 12 * m1, 1e9, "GB/s", b,
                               "saxpy n=2^26 incx 2 incy -3, 1 thread"),
        "sdot_inc2_-3": _port(lambda: o.dot(m1, x, 2, y, -3), 8 * m1, 1e9, "GB/s", b,
                              "sdot n=2^26 incx 2 incy -3, 1 thread"),
        "sgemv_N_inc2_-3": _port(lambda: o.gemv(111, m2, m2, 1.5, 0.5, A, m2, xv, 2, yv, -3),
                                 (m2 * m2 + 3 * m2) * 4, 1e9, "GB/s", b, "sgemv N 8192^2 incx 2 incy -3, 1 thread"),
    }


CPU_LEGS = {"bandwidth": cpu_bandwidth, "config1": cpu_config1, "config3": cpu_config3, "config4": cpu_config4,
            "config5": cpu_config5, "families": cpu_families, "strided": cpu_strided}

CONFIGS = {"config1": config1, "config3": config3, "config4": config4, "config5": config5, "families": families,
           "strided": strided}


def run_config(name):
    def w(a, torch, so, L):
        r = CONFIGS[name](a, torch, so, L)
        roof = r.get("roofline")
        return _line(f"{name} of the default run", None, None, a, 0.0, "see kernels", r["workload"], roof,
                     r.get("kernels", {}), section=r)
    return w


def bench_default(a, torch, so, L):
    """The bandwidth-set headline line, with every other configuration and
    the remaining kernel families under "configs" (skipped with
    --no-configs)."""
    out = bench_bandwidth(a, torch, so, L)
    if not a.no_configs:
        out["configs"] = {}
        for name, fn in CONFIGS.items():
            t0 = time.perf_counter()
            print(f"bench: {name} ...", file=sys.stderr, flush=True)
            r = fn(a, torch, so, L)
            r["wall_s"] = round(time.perf_counter() - t0, 1)
            out["configs"][name] = r
    return out


WORKLOADS = {"default": bench_default, "bandwidth": bench_bandwidth, "level3": bench_level3, "trsv": bench_trsv,
             "extended": bench_extended, "precision": bench_precision, "serving": bench_serving,
             "hostpath": bench_hostpath}
WORKLOADS.update({name: run_config(name) for name in CONFIGS})


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="default")
    ap.add_argument("--no-configs", action="store_true",
                    help="default workload: the bandwidth set only, without the other configurations")
    ap.add_argument("--n1", type=int, default=1 << 26)
    ap.add_argument("--n2", type=int, default=8192)
    ap.add_argument("--n-gemm", type=int, default=8192)
    ap.add_argument("--requests", type=int, default=20000)
    ap.add_argument("--rate", type=float, default=20000.0)
    ap.add_argument("--max-batch", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--gbench-dir", default=None,
                    help="also write the measured runs as Google Benchmark JSON (<dir>/exoblas_amd/<kernel>.json)")
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    if a.cpu_leg:  # a child process: CPU baselines only, no torch, no GPU
        print(json.dumps(CPU_LEGS[a.cpu_leg](a)), flush=True)
        return 0
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2
    import torch

    from exoblas_amd import lib as L
    so = L.load()
    torch.cuda.set_device(0)
    out = WORKLOADS[a.workload](a, torch, so, L)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
