"""Google-Benchmark-compatible JSON for runs measured by bench.py.

The reference's benchmark programs register each run under the name built
by kernel_name() (test/common/misc.h:97-124): "<lib>_<p><kernel>", then
"_rm" for row-major, then "_" and the Side / Uplo / Trans / Diag letters
when any of those is set (misc.h:46-95), followed by Google Benchmark's
"/<ArgName>:<value>" pairs in the order of the bench.cpp's ArgNames (e.g.
test/level1/axpy/bench.cpp:39-40, test/level2/gemv/bench.cpp).  Its plotting
tools read one JSON file per kernel from <dir>/<library>/<kernel>.json
(analytics_tools/graphing/graph.py:110-135), take "real_time" in ns, and
split the run name at "/" and ":" (analytics_tools/graphing/misc.py:24-30),
so records written here in the same shape plot beside the CPU libraries'.
"""
from __future__ import annotations

import json
import os
import platform
import time

ROW_MAJOR, COL_MAJOR = 101, 102
NO_TRANS, TRANS = 111, 112
UPPER, LOWER = 121, 122
NON_UNIT, UNIT = 131, 132
LEFT, RIGHT = 141, 142

# bench type enum of test/common/bench_ranges.h:18-25
LEVEL_1, LEVEL_2_EQ, LEVEL_2_SQ, LEVEL_3_EQ = 0, 1, 2, 3


def _order(o):
    return {ROW_MAJOR: "_rm", COL_MAJOR: "_col"}.get(o, "")


def _side(s):
    return {LEFT: "l", RIGHT: "r"}.get(s, "")


def _trans(t):
    return {NO_TRANS: "n", TRANS: "t"}.get(t, "")


def _uplo(u):
    return {LOWER: "l", UPPER: "u"}.get(u, "")


def _diag(d):
    return {NON_UNIT: "n", UNIT: "u"}.get(d, "")


def kernel_name(lib: str, p: str, kernel: str, order=None, side=0, uplo=0, trans_a=0, trans_b=0, diag=0,
                level3=False) -> str:
    """misc.h:97-124: the level-1 form when order is None, else the level-2
    (Order, Uplo, TransA, Diag) or level-3 (Order, Side, Uplo, TransA,
    TransB, Diag) form."""
    name = f"{lib}_{p}{kernel}"
    if order is None:
        return name
    name += _order(order)
    if level3:
        if side or uplo or trans_a or trans_b:
            name += "_"
        return name + _side(side) + _uplo(uplo) + _trans(trans_a) + _trans(trans_b) + _diag(diag)
    if uplo or trans_a or diag:
        name += "_"
    return name + _uplo(uplo) + _trans(trans_a) + _diag(diag)


def run_name(name: str, args: list[tuple[str, object]]) -> str:
    return name + "".join(f"/{k}:{v}" for k, v in args)


def record(name: str, args: list[tuple[str, object]], real_time_ns: float, iterations: int, **counters) -> dict:
    rn = run_name(name, args)
    rec = {"name": rn, "run_name": rn, "run_type": "iteration", "repetitions": 1, "repetition_index": 0,
           "threads": 1, "iterations": int(iterations), "real_time": float(real_time_ns),
           "cpu_time": float(real_time_ns), "time_unit": "ns"}
    rec.update({k: float(v) for k, v in counters.items()})
    return rec


def write(root: str, lib_dir: str, kernel: str, records: list[dict]) -> str:
    """<root>/<lib_dir>/<kernel>.json in Google Benchmark's layout."""
    d = os.path.join(root, lib_dir)
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"{kernel}.json")
    doc = {"context": {"date": time.strftime("%Y-%m-%dT%H:%M:%S%z"), "host_name": platform.node(),
                       "executable": "bench.py (exoblas_amd, MI355X, device-resident operands)",
                       "num_cpus": os.cpu_count() or 1, "mhz_per_cpu": 0, "cpu_scaling_enabled": False,
                       "caches": [], "library_build_type": "release"},
           "benchmarks": records}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path


def _bits(p):
    return 32 if p == "s" else 64


def level1_records(p: str, kind: str, n: int, us: float, iterations: int, alpha=0.75, c=0.6, s=0.8) -> tuple[str, dict]:
    """(kernel file, record) for a unit-stride level-1 run of n elements
    (ArgNames of test/level1/<kernel>/bench.cpp)."""
    al = [("alignmentX", 64), ("alignmentY", 64), ("bench_type", LEVEL_1), ("precision", _bits(p))]
    if kind == "axpy":
        args = [("N", n), ("alpha", alpha), ("incX", 1), ("incY", 1)] + al
    elif kind == "rot":
        args = [("N", n), ("incX", 1), ("incY", 1), ("c", c), ("s", s)] + al
    else:  # dot, copy, swap
        args = [("N", n), ("incX", 1), ("incY", 1)] + al
    return kind, record(kernel_name("exo", p, kind), args, us * 1e3, iterations)


def gemv_record(p: str, trans: int, m: int, n: int, us: float, iterations: int, alpha=1.5, beta=0.5) -> dict:
    """test/level2/gemv/bench.cpp ArgNames; square problems are level_2_sq."""
    args = [("M", m), ("N", n), ("Order", ROW_MAJOR), ("TransA", trans), ("alpha", alpha), ("lda_diff", 0),
            ("incX", 1), ("beta", beta), ("incY", 1), ("alignmentA", 64), ("alignmentX", 64), ("alignmentY", 64),
            ("bench_type", LEVEL_2_SQ if m == n else LEVEL_2_EQ), ("precision", _bits(p))]
    return record(kernel_name("exo", p, "gemv", ROW_MAJOR, trans_a=trans), args, us * 1e3, iterations)


def trsv_record(p: str, uplo: int, trans: int, diag: int, n: int, us: float, iterations: int) -> dict:
    """test/level2/trsv/bench.cpp ArgNames."""
    args = [("N", n), ("order", ROW_MAJOR), ("Uplo", uplo), ("TransA", trans), ("Diag", diag), ("lda_diff", 0),
            ("incX", 1), ("alignmentA", 64), ("alignmentX", 64), ("bench_type", LEVEL_2_SQ), ("precision", _bits(p))]
    return record(kernel_name("exo", p, "trsv", ROW_MAJOR, uplo=uplo, trans_a=trans, diag=diag), args, us * 1e3, iterations)


def gemm_record(p: str, ta: int, tb: int, m: int, n: int, k: int, us: float, iterations: int, alpha=1.5,
                beta=0.0) -> dict:
    """test/level3/gemm/bench.cpp ArgNames."""
    args = [("M", m), ("N", n), ("K", k), ("order", ROW_MAJOR), ("TransA", ta), ("TransB", tb), ("alpha", alpha),
            ("lda_diff", 0), ("ldb_diff", 0), ("beta", beta), ("ldc_diff", 0), ("alignmentA", 64),
            ("alignmentB", 64), ("alignmentC", 64), ("bench_type", LEVEL_3_EQ), ("precision", _bits(p))]
    return record(kernel_name("exo", p, "gemm", ROW_MAJOR, trans_a=ta, trans_b=tb, level3=True), args, us * 1e3,
                  iterations)


def syrk_record(p: str, uplo: int, trans: int, n: int, k: int, us: float, iterations: int, alpha=1.5,
                beta=0.0) -> dict:
    """test/level3/syrk/bench.cpp ArgNames (registered with Side 0, TransB 0)."""
    args = [("N", n), ("K", k), ("order", ROW_MAJOR), ("Uplo", uplo), ("Trans", trans), ("alpha", alpha),
            ("lda_diff", 0), ("beta", beta), ("ldc_diff", 0), ("alignmentA", 64), ("alignmentC", 64),
            ("bench_type", LEVEL_3_EQ), ("precision", _bits(p))]
    return record(kernel_name("exo", p, "syrk", ROW_MAJOR, uplo=uplo, trans_a=trans, level3=True), args, us * 1e3,
                  iterations)


def symm_record(p: str, side: int, uplo: int, m: int, n: int, us: float, iterations: int, alpha=1.5,
                beta=0.0) -> dict:
    """test/level3/symm/bench.cpp ArgNames."""
    args = [("M", m), ("N", n), ("Order", ROW_MAJOR), ("Side", side), ("Uplo", uplo), ("alpha", alpha),
            ("lda_diff", 0), ("ldb_diff", 0), ("beta", beta), ("ldc_diff", 0), ("alignmentA", 64),
            ("alignmentB", 64), ("alignmentC", 64), ("bench_type", LEVEL_3_EQ), ("precision", _bits(p))]
    return record(kernel_name("exo", p, "symm", ROW_MAJOR, side=side, uplo=uplo, level3=True), args, us * 1e3,
                  iterations)


def syr2k_record(p: str, uplo: int, trans: int, n: int, k: int, us: float, iterations: int, alpha=17,
                 beta=18) -> dict:
    """test/level3/syr2k/bench.cpp:43-64 ArgNames (the same kernel_name scheme as syrk)."""
    args = [("N", n), ("K", k), ("order", ROW_MAJOR), ("Uplo", uplo), ("Trans", trans), ("alpha", alpha),
            ("lda_diff", 0), ("ldb_diff", 0), ("beta", beta), ("ldc_diff", 0), ("alignmentA", 64),
            ("alignmentB", 64), ("alignmentC", 64), ("bench_type", LEVEL_3_EQ), ("precision", _bits(p))]
    return record(kernel_name("exo", p, "syr2k", ROW_MAJOR, uplo=uplo, trans_a=trans, level3=True), args, us * 1e3,
                  iterations)


def strmm_record(uplo: int, trans: int, diag: int, m: int, n: int, alpha: int, us: float, iterations: int) -> dict:
    """test/level3/strmm/bench.cpp:103-109: the benchmark BM_exo_strmm, enums encoded 0/1 (Order 0 = row
    major, Side 0 = Left, Uplo 0 = Lower, TransA 0 = NoTrans, Diag 0 = NonUnit), alpha 0 or 1."""
    args = [("m", m), ("n", n), ("Order", 0), ("Side", 0), ("Uplo", 0 if uplo == 122 else 1),
            ("TransA", 0 if trans == 111 else 1), ("Diag", 0 if diag == 131 else 1), ("alpha", alpha),
            ("lda", m), ("ldb", n), ("alignmentA", 64), ("alignmentB", 64)]
    return record("BM_exo_strmm", args, us * 1e3, iterations)
