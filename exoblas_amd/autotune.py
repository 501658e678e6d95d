"""Autotuner for the batched kernel's gemm tile choice.

For every shape class of the tuning table (M class x N class x K class, the
classes of include/exoblas_amd_batch.h) it times a seeded batch of gemms drawn
from that class with each tile forced (32, 64, 128), takes the median of
`reps` launches, and picks the fastest tile; a smaller tile within `tie` of
the fastest wins, so the choice is stable from run to run.  The table and the
timings go to exoblas_amd/tuning.json; ``apply()`` loads the table into the
library (the serving loop calls it).

    python -m exoblas_amd.autotune [--reps 5] [--batch 48]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys

import numpy as np

from . import lib as L
from . import native as N

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuning.json")
M_EDGES = [1, 32, 64, 128, 256, 512, 1024]   # class c covers (M_EDGES[c], M_EDGES[c+1]]
K_EDGES = [1, 256, 1024]
TILES = (32, 64, 128)


def apply(path: str = TABLE_PATH) -> bool:
    """Loads a tuned table into the library; False if there is none."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        table = json.load(f)["table"]
    arr = (ctypes.c_int32 * 72)(*table)
    if L.load().exoblas_batch_set_tuning(arr, 72) != 0:
        raise RuntimeError(L.last_error())
    return True


def _class_batch(torch, rng, mc, nc, kc, count, pools):
    lo_m, hi_m = M_EDGES[mc] + 1, M_EDGES[mc + 1]
    lo_n, hi_n = M_EDGES[nc] + 1, M_EDGES[nc + 1]
    lo_k, hi_k = K_EDGES[kc] + 1, K_EDGES[kc + 1]
    gA, gB, out = pools
    ops = np.zeros(count, N.OP_DTYPE)
    ops["kind"] = N.OP_GEMM
    ops["trans_a"] = ops["trans_b"] = 111
    ops["m"] = rng.integers(max(1, lo_m), hi_m + 1, count)
    ops["n"] = rng.integers(max(1, lo_n), hi_n + 1, count)
    ops["k"] = rng.integers(max(1, lo_k), hi_k + 1, count)
    ops["a"], ops["lda"] = gA.data_ptr(), 1024
    ops["b"], ops["ldb"] = gB.data_ptr(), 1024
    off = np.concatenate([[0], np.cumsum(ops["m"] * ops["n"] + 64)])[:-1]
    assert off[-1] + ops["m"][-1] * ops["n"][-1] <= out.numel()
    ops["c"] = out.data_ptr() + 4 * off
    ops["ldc"] = ops["n"]
    ops["alpha"], ops["beta"] = 1.0, 0.5
    return ops


def tune(reps: int = 5, batch: int = 48, tie: float = 0.03, seed: int = 0):
    import torch
    so = L.load()
    stream = torch.cuda.current_stream()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    gA = torch.empty((1024, 1024), device="cuda").uniform_(-1, 1, generator=g)
    gB = torch.empty((1024, 1024), device="cuda").uniform_(-1, 1, generator=g)
    out = torch.zeros(batch * (1024 * 1024 + 64), device="cuda")
    rng = np.random.default_rng(seed)
    table = [0] * 72
    times = {}
    try:
        for kc in range(2):
            for mc in range(6):
                for nc in range(6):
                    ops = _class_batch(torch, rng, mc, nc, kc, batch, (gA, gB, out))
                    t = []
                    for tile in range(3):
                        so.exoblas_batch_force_tile(tile)
                        so.exoblas_batch_run(stream.cuda_stream, ops.ctypes.data, len(ops))  # warm-up
                        ms = []
                        for _ in range(reps):
                            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                            e0.record(stream)
                            if so.exoblas_batch_run(stream.cuda_stream, ops.ctypes.data, len(ops)) != 0:
                                raise RuntimeError(L.last_error())
                            e1.record(stream)
                            e1.synchronize()
                            ms.append(e0.elapsed_time(e1))
                        t.append(float(np.median(ms)))
                    best = min(t)
                    choice = next(i for i in range(3) if t[i] <= best * (1 + tie))
                    table[kc * 36 + mc * 6 + nc] = choice
                    times[f"k{kc}_m{mc}_n{nc}"] = [round(x, 4) for x in t]
    finally:
        so.exoblas_batch_force_tile(-1)
    return table, times


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--batch", type=int, default=48)
    ap.add_argument("--out", default=TABLE_PATH)
    a = ap.parse_args(argv)
    table, times = tune(a.reps, a.batch)
    with open(a.out, "w") as f:
        json.dump({"tiles": list(TILES), "m_edges": M_EDGES, "k_edges": K_EDGES, "table": table,
                   "median_ms_per_tile": times, "batch": a.batch, "reps": a.reps}, f, indent=1)
    print(json.dumps({"table": table}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
