"""The C-ABI boundary, checked without a GPU: the library loads, exports every
function the public headers declare, the window structs have the reference's
layout, and the generated files are current."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

_DECL = re.compile(r"^\s*(?:void|int|const char \*|float|double)\s*\**\s*(\w+)\s*\(", re.M)


def declared_functions():
    names = set()
    for f in sorted(os.listdir(INC)):
        if f.endswith(".h"):
            with open(os.path.join(INC, f)) as fh:
                names.update(_DECL.findall(fh.read()))
    return names


def test_library_exports_every_declared_symbol():
    from exoblas_amd import lib as L
    so = ctypes.CDLL(L.LIB_PATH)
    names = declared_functions()
    assert len(names) > 200
    missing = [n for n in sorted(names) if not hasattr(so, n)]
    assert not missing, missing


def test_reference_symbol_set():
    """The exo_* names the reference wrappers call (test/level*/<k>/exo_<k>_wrapper.h)."""
    from exoblas_amd import generator as G
    syms = set(G.all_symbols())
    for s in ["exo_saxpy_stride_1", "exo_daxpy_alpha_1_stride_1", "exo_sdot_stride_any", "exo_dsdot_stride_1",
              "exo_sdsdot_stride_any", "exo_srotm_flag_neg_one_stride_1", "exo_drotm_flag_one_stride_any",
              "exo_sscal_alpha_0_stride_1", "exo_sgemv_rm_stride_1", "exo_dtrsv_rm_stride_any",
              "exo_ssyr2_rm_stride_1", "exo_sgemm_stride_1", "exo_smscal_rm_stride_1",
              "exo_dtrmscal_rm_stride_1", "exo_ssymm_rm_stride_1", "exo_dsyrk_rm_stride_1"]:
        assert s in syms, s


def test_window_struct_layout():
    src = r'''
#include <stddef.h>
#include "exo_types.h"
_Static_assert(sizeof(int_fast32_t) == 8, "int_fast32_t");
_Static_assert(sizeof(struct exo_win_1f32) == 16, "w1");
_Static_assert(sizeof(struct exo_win_1f64c) == 16, "w1c");
_Static_assert(sizeof(struct exo_win_2f32c) == 24, "w2");
_Static_assert(offsetof(struct exo_win_2f64, strides) == 8, "off");
int main(void) { return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.c")
        with open(p, "w") as f:
            f.write(src)
        r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-I", INC, p], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    from exoblas_amd import lib as L
    assert ctypes.sizeof(L.Win1) == 16 and ctypes.sizeof(L.Win2) == 24


def test_headers_compile_as_c_and_cpp():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.c")
        with open(p, "w") as f:
            for h in sorted(os.listdir(INC)):
                if h.startswith("exo_") and h.endswith(".h"):
                    f.write(f'#include "{h}"\n')
            f.write('#include "exoblas_amd.h"\nint main(void){return 0;}\n')
        for cc in (["gcc", "-std=c11"], ["g++", "-x", "c++", "-std=c++17"]):
            r = subprocess.run(cc + ["-fsyntax-only", "-I", INC, p], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr


def test_generated_files_are_current():
    from exoblas_amd import generator as G
    for path, text in G.outputs().items():
        with open(path) as f:
            assert f.read() == text, f"{path} is stale: run python -m exoblas_amd.generator"


def test_no_cpu_fallback_in_product():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "exoblas_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".cpp", ".hip", ".hpp")):
                with open(os.path.join(dp, f)) as fh:
                    txt = fh.read()
                assert "import oracle" not in txt and "from oracle" not in txt and "liboracle" not in txt, f


def test_cmake_targets_match_the_generated_libraries():
    """exoblas_amd/cmake/EXO_BLASConfig.cmake defines one EXO_BLAS::exo_<lib>
    target per generated header (the reference's library targets)."""
    import re

    from exoblas_amd import generator as G
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "exoblas_amd", "cmake", "EXO_BLASConfig.cmake")).read()
    m = re.search(r"foreach\(_k ([^)]*)\)", text)
    assert m, "no target list"
    assert sorted(m.group(1).split()) == sorted(G.LIB_ORDER)
    assert len(G.LIB_ORDER) == 26
