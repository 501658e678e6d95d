"""Drop-in check without a GPU: the reference's own CBLAS-style wrapper
headers (test/level{1,2,3}/<k>/exo_<k>_wrapper.h) compile against this
build's include/ directory and link against libexoblas_amd.so with every exo_*
symbol they call resolved.  The headers are read from the reference checkout
when it is present (they are not copied into this repository); the test is
skipped where it is absent.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

WRAPPERS = {
    "level1": [("asum", "exo_asum_wrapper.h", "exo_sasum"), ("axpy", "exo_axpy_wrapper.h", "exo_daxpy"),
               ("copy", "exo_copy_wrapper.h", "exo_scopy"), ("dot", "exo_dot_wrapper.h", "exo_ddot"),
               ("dsdot", "exo_dsdot_wrapper.h", "exo_dsdot"), ("rot", "exo_rot_wrapper.h", "exo_srot"),
               ("rotm", "exo_rotm_wrapper.h", "exo_drotm"), ("scal", "exo_scal_wrapper.h", "exo_sscal"),
               ("sdsdot", "exo_sdsdot.h", "exo_sdsdot"), ("swap", "exo_swap_wrapper.h", "exo_dswap")],
    "level2": [("gemv", "exo_gemv_wrapper.h", "exo_sgemv"), ("ger", "exo_ger_wrapper.h", "exo_dger"),
               ("symv", "exo_symv_wrapper.h", "exo_ssymv"), ("syr", "exo_syr_wrapper.h", "exo_dsyr"),
               ("syr2", "exo_syr2_wrapper.h", "exo_ssyr2"), ("trmv", "exo_trmv_wrapper.h", "exo_dtrmv"),
               ("trsv", "exo_trsv_wrapper.h", "exo_strsv"), ("sgbmv", "exo_sgbmv.h", "exo_sgbmv"),
               ("dgbmv", "exo_dgbmv.h", "exo_dgbmv"), ("ssbmv", "exo_ssbmv.h", "exo_ssbmv"),
               ("dsbmv", "exo_dsbmv.h", "exo_dsbmv"), ("stbmv", "exo_stbmv.h", "exo_stbmv"),
               ("dtbmv", "exo_dtbmv.h", "exo_dtbmv"), ("stbsv", "exo_stbsv.h", "exo_stbsv"),
               ("dtbsv", "exo_dtbsv.h", "exo_dtbsv")],
    "level3": [("gemm", "exo_gemm_wrapper.h", "exo_sgemm"), ("symm", "exo_symm_wrapper.h", "exo_dsymm"),
               ("syrk", "exo_syrk_wrapper.h", "exo_ssyrk"), ("syr2k", "exo_syr2k_wrapper.h", "exo_dsyr2k"),
               ("strmm", "exo_strmm.h", "exo_strmm")],
}

CASES = [(lvl, k, h, fn) for lvl, v in WRAPPERS.items() for (k, h, fn) in v]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "test", "common")), reason="reference checkout absent")
@pytest.mark.parametrize("lvl,kernel,header,fn", CASES)
def test_reference_wrapper_links(lvl, kernel, header, fn):
    lib = os.path.join(ROOT, "exoblas_amd", "libexoblas_amd.so")
    assert os.path.exists(lib), "build the library first"
    wdir = os.path.join(REF, "test", lvl, kernel)
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        with open(src, "w") as f:
            f.write(f'#include <cblas.h>\n#include "{header}"\n#include <cstdio>\n'
                    f'int main() {{ std::printf("%p\\n", (void*)&{fn}); return 0; }}\n')
        exe = os.path.join(d, "t")
        cmd = ["g++", "-std=c++20", "-w", "-I", wdir, "-I", os.path.join(REF, "test", "common"),
               "-I", os.path.join(ROOT, "include"), src, "-o", exe, lib, "-Wl,--no-undefined",
               "-Wl,--unresolved-symbols=report-all", f"-Wl,-rpath,{os.path.dirname(lib)}",
               "-Wl,--allow-shlib-undefined"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
