"""Google-Benchmark-compatible output: run names follow the reference's
kernel_name() rules (test/common/misc.h:46-124) and the records carry what
analytics_tools/graphing reads (kernels_specs.py:6-30, misc.py:24-30)."""
from __future__ import annotations

import json

import pytest

from exoblas_amd import gbench as g


@pytest.mark.parametrize("name,expect", [
    (g.kernel_name("exo", "s", "axpy"), "exo_saxpy"),
    (g.kernel_name("exo", "s", "gemv", g.ROW_MAJOR, trans_a=g.NO_TRANS), "exo_sgemv_rm_n"),
    (g.kernel_name("exo", "d", "gemv", g.ROW_MAJOR, trans_a=g.TRANS), "exo_dgemv_rm_t"),
    (g.kernel_name("exo", "s", "trsv", g.ROW_MAJOR, uplo=g.LOWER, trans_a=g.NO_TRANS, diag=g.NON_UNIT),
     "exo_strsv_rm_lnn"),
    (g.kernel_name("exo", "s", "trsv", g.ROW_MAJOR, uplo=g.UPPER, trans_a=g.TRANS, diag=g.UNIT), "exo_strsv_rm_utu"),
    (g.kernel_name("exo", "s", "ger", g.ROW_MAJOR), "exo_sger_rm"),
    (g.kernel_name("exo", "s", "gemm", g.ROW_MAJOR, trans_a=g.NO_TRANS, trans_b=g.TRANS, level3=True),
     "exo_sgemm_rm_nt"),
    (g.kernel_name("exo", "s", "syrk", g.ROW_MAJOR, uplo=g.LOWER, trans_a=g.NO_TRANS, level3=True), "exo_ssyrk_rm_ln"),
    (g.kernel_name("exo", "d", "symm", g.ROW_MAJOR, side=g.LEFT, uplo=g.LOWER, level3=True), "exo_dsymm_rm_ll"),
])
def test_kernel_names(name, expect):
    assert name == expect


def _parse(run_name):
    """The graphing tools' split of a run name into fields."""
    toks = [t.split(":") for t in ("sub_kernel_name:" + run_name).split("/")]
    return {k: v for k, v in toks}


def test_records_and_files(tmp_path):
    recs = [g.gemv_record("s", g.NO_TRANS, 8192, 8192, 44.7, 10), g.trsv_record("d", g.LOWER, g.NO_TRANS,
                                                                                  g.NON_UNIT, 4096, 298.0, 20)]
    kind, r1 = g.level1_records("d", "axpy", 1 << 26, 250.0, 10)
    assert kind == "axpy"
    for r in recs + [r1]:
        assert r["name"] == r["run_name"] and r["time_unit"] == "ns" and r["real_time"] > 0
        d = _parse(r["run_name"])
        assert d["sub_kernel_name"].startswith("exo_")
        assert int(d["bench_type"]) in (0, 1, 2, 3) and d["precision"] in ("32", "64")
    d = _parse(recs[0]["run_name"])
    assert (int(d["M"]), int(d["N"]), int(d["Order"]), int(d["TransA"])) == (8192, 8192, 101, 111)
    assert recs[0]["real_time"] == pytest.approx(44700.0)
    path = g.write(str(tmp_path), "exoblas_amd", "gemv", recs)
    doc = json.load(open(path))
    assert doc["benchmarks"][0]["run_name"].startswith("exo_sgemv_rm_n/M:8192/N:8192")
    assert path.endswith("exoblas_amd/gemv.json")


def test_syr2k_and_strmm_records():
    r = g.syr2k_record("s", g.LOWER, g.NO_TRANS, 4096, 1024, 390.2, 5, alpha=1.5, beta=1)
    d = _parse(r["run_name"])
    assert d["sub_kernel_name"] == "exo_ssyr2k_rm_ln"
    assert (int(d["N"]), int(d["K"]), int(d["Uplo"]), int(d["Trans"])) == (4096, 1024, g.LOWER, g.NO_TRANS)
    assert int(d["bench_type"]) == g.LEVEL_3_EQ and d["precision"] == "32"
    t = g.strmm_record(g.LOWER, g.NO_TRANS, g.NON_UNIT, 4096, 4096, 1, 1070.0, 5)
    # test/level3/strmm/bench.cpp:107-109: BM_exo_strmm with enums as 0/1 (0 = Lower / NoTrans / NonUnit)
    assert t["run_name"] == ("BM_exo_strmm/m:4096/n:4096/Order:0/Side:0/Uplo:0/TransA:0/Diag:0/alpha:1/"
                             "lda:4096/ldb:4096/alignmentA:64/alignmentB:64")
    assert t["real_time"] == pytest.approx(1070000.0)
