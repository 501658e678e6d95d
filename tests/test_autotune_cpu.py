"""The autotuner's table format and loading path, without a GPU: a table
written in the tuner's JSON format is loaded into the library and read back;
the committed table (when present) has 72 valid tile ids."""
from __future__ import annotations

import ctypes
import json
import os

import pytest


def test_apply_roundtrip(tmp_path):
    from exoblas_amd import autotune, lib
    so = lib.load()
    saved = (ctypes.c_int32 * 72)()
    assert so.exoblas_batch_get_tuning(saved, 72) == 0
    table = [(i * 7) % 3 for i in range(72)]
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"table": table}))
    try:
        assert autotune.apply(str(p))
        out = (ctypes.c_int32 * 72)()
        so.exoblas_batch_get_tuning(out, 72)
        assert list(out) == table
    finally:
        so.exoblas_batch_set_tuning(saved, 72)
    assert not autotune.apply(str(tmp_path / "missing.json"))


def test_committed_table_valid():
    from exoblas_amd import autotune
    if not os.path.exists(autotune.TABLE_PATH):
        pytest.skip("no tuned table committed yet")
    with open(autotune.TABLE_PATH) as f:
        d = json.load(f)
    assert len(d["table"]) == 72 and all(v in (0, 1, 2) for v in d["table"])
    assert d["tiles"] == [32, 64, 128]
