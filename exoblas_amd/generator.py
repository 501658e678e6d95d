"""Variant generator: the exported exo_* symbol set, its headers and C-ABI glue.

This module takes the place of the reference's build-time scheduling layer
(src/common/codegen_helpers.py:188-233 ``variants_generator``,
src/common/blaslib.py:25-56 ``optimize_level_1`` and the per-kernel schedules
in src/level{1,2,3}/*.py).  Where the reference rewrites each ``@proc`` into a
vectorised C loop nest per precision, this generator maps each
(kernel, precision, stride variant) onto a hand-written HIP kernel family for
gfx950 and emits:

* ``include/exo_<lib>.h`` -- one header per reference library target, with the
  same function names and argument lists as the Exo-generated headers (the
  names are pinned by the reference wrapper call sites, cited per proc below);
* ``include/exoblas_amd_async.h`` -- the device-pointer/stream API: the same
  argument lists with ``hipStream_t`` in place of ``void *ctxt``;
* ``exoblas_amd/csrc/gen_exports.cpp`` -- the extern "C" definitions.  Each
  validates enum arguments against the CBLAS domain (the reference asserts
  them, codegen_helpers.py:87-90), stages host windows to device memory when
  needed, and calls the adapter ``exoblas::ops::<proc><T>`` with plain device
  pointers, element strides and scalar values.

Tile shapes and launch parameters are not chosen here: the kernel adapters
pick them (``csrc/level3.hip`` plans, ``csrc/batch.hip`` tile classes), and the
batched path reads the autotuned table ``exoblas_amd/tuning.json`` written by
``exoblas_amd/autotune.py``.

Run ``python -m exoblas_amd.generator`` to rewrite the files; the test suite
checks that the committed files equal what this module produces.
"""
from __future__ import annotations

import dataclasses
import os
import sys
from typing import List, Optional, Sequence, Tuple

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

NO_TRANS = 111
TRANS_VALUES = (111, 112, 113)
UPLO_VALUES = (121, 122)
DIAG_VALUES = (131, 132)
SIDE_VALUES = (141, 142)

PRECISIONS = {"s": ("float", "f32"), "d": ("double", "f64")}


@dataclasses.dataclass
class Arg:
    kind: str  # size | enum | scalar | result | win | array
    name: str
    dims: Tuple[str, ...] = ()
    const: bool = True
    values: Tuple[int, ...] = ()
    used_when: Optional[str] = None  # C expression over size/enum args
    length: int = 0  # array length
    ctype: Optional[str] = None  # overrides the precision's element type


def size(n):
    return Arg("size", n)


def enum(n, values):
    return Arg("enum", n, values=tuple(values))


def scalar(n, ctype=None):
    return Arg("scalar", n, ctype=ctype)


def result(n, ctype=None):
    return Arg("result", n, const=False, ctype=ctype)


def win(n, *dims, const=True, used_when=None, ctype=None):
    return Arg("win", n, dims=tuple(dims), const=const, used_when=used_when, ctype=ctype)


def array(n, length):
    return Arg("array", n, length=length)


@dataclasses.dataclass
class Proc:
    name: str  # reference proc name, e.g. "axpy" or "gemv_rm"
    lib: str  # header / library target, e.g. "axpy" -> exo_axpy.h
    args: List[Arg]
    ref: str  # reference spec citation
    callers: str  # reference wrapper call sites
    precisions: Sequence[str] = ("s", "d")
    fixed_prefix: Optional[str] = None  # dsdot/sdsdot carry no s/d prefix
    # name suffixes the reference exports; "" = the bare name (trmm's procs are
    # exported without the stride_1 / stride_any pair)
    variants: Sequence[str] = ("stride_any", "stride_1")


W = "int_fast32_t"

PROCS: List[Proc] = [
    # ------------------------------------------------------------ level 1
    Proc("asum", "asum", [size("n"), win("x", "n"), result("result")],
         "src/level1/asum.py:11-15", "test/level1/asum/exo_asum_wrapper.h:9-20"),
    Proc("axpy", "axpy", [size("n"), scalar("alpha"), win("x", "n"), win("y", "n", const=False)],
         "src/level1/axpy.py:9-12", "test/level1/axpy/exo_axpy_wrapper.h:17-32"),
    Proc("axpy_alpha_1", "axpy", [size("n"), win("x", "n"), win("y", "n", const=False)],
         "src/level1/axpy.py:15-18", "test/level1/axpy/exo_axpy_wrapper.h:13-15"),
    Proc("dot", "dot", [size("n"), win("x", "n"), win("y", "n"), result("result")],
         "src/level1/dot.py:9-13", "test/level1/dot/exo_dot_wrapper.h:10-24"),
    Proc("dsdot", "dsdot",
         [size("n"), win("x", "n", ctype="float"), win("y", "n", ctype="float"),
          result("result", ctype="double")],
         "src/level1/dsdot.py:11-21", "test/level1/dsdot/exo_dsdot_wrapper.h:9-22",
         precisions=("s",), fixed_prefix=""),
    Proc("sdsdot", "dsdot",
         [size("n"), scalar("sb", ctype="float"), win("x", "n", ctype="float"),
          win("y", "n", ctype="float"), result("result", ctype="float")],
         "src/level1/dsdot.py:24-36", "test/level1/sdsdot/exo_sdsdot.h:9-23",
         precisions=("s",), fixed_prefix=""),
    Proc("copy", "copy", [size("n"), win("x", "n"), win("y", "n", const=False)],
         "src/level1/exo_copy.py:9-12", "test/level1/copy/exo_copy_wrapper.h:9-21"),
    Proc("swap", "swap", [size("n"), win("x", "n", const=False), win("y", "n", const=False)],
         "src/level1/swap.py:9-15", "test/level1/swap/exo_swap_wrapper.h:9-21"),
    Proc("rot", "rot",
         [size("n"), win("x", "n", const=False), win("y", "n", const=False), scalar("c"), scalar("s")],
         "src/level1/rot.py:9-15", "test/level1/rot/exo_rot_wrapper.h:9-21"),
    Proc("rotm_flag_neg_one", "rotm",
         [size("n"), win("x", "n", const=False), win("y", "n", const=False), array("H", 4)],
         "src/level1/rotm.py:9-15", "test/level1/rotm/exo_rotm_wrapper.h:10-13,30-33"),
    Proc("rotm_flag_zero", "rotm",
         [size("n"), win("x", "n", const=False), win("y", "n", const=False), array("H", 4)],
         "src/level1/rotm.py:18-24", "test/level1/rotm/exo_rotm_wrapper.h:14-17,34-37"),
    Proc("rotm_flag_one", "rotm",
         [size("n"), win("x", "n", const=False), win("y", "n", const=False), array("H", 4)],
         "src/level1/rotm.py:27-33", "test/level1/rotm/exo_rotm_wrapper.h:18-21,38-41"),
    Proc("scal", "scal", [size("n"), scalar("alpha"), win("x", "n", const=False)],
         "src/level1/scal.py:9-12", "test/level1/scal/exo_scal_wrapper.h:18-20,30-32"),
    Proc("scal_alpha_0", "scal", [size("n"), win("x", "n", const=False)],
         "src/level1/scal.py:15-19", "test/level1/scal/exo_scal_wrapper.h:15-16,27-28"),
    # ------------------------------------------------------------ level 2
    Proc("gemv_rm", "gemv",
         [enum("TransA", TRANS_VALUES), size("M"), size("N"), scalar("alpha"), scalar("beta"),
          win("A", "M", "N", used_when="TransA == 111"),
          win("AT", "N", "M", used_when="TransA != 111"),
          win("x", "N"), win("y", "M", const=False)],
         "src/level2/gemv.py:10-22", "test/level2/gemv/exo_gemv_wrapper.h:19-37"),
    Proc("ger_rm", "ger",
         [size("m"), size("n"), scalar("alpha"), win("x", "m"), win("y", "n"),
          win("A", "m", "n", const=False)],
         "src/level2/ger.py:9-15", "test/level2/ger/exo_ger_wrapper.h:16-32"),
    Proc("symv_rm", "symv",
         [enum("Uplo", UPLO_VALUES), size("n"), scalar("alpha"), win("A", "n", "n"), win("x", "n"),
          scalar("beta"), win("y", "n", const=False)],
         "src/level2/symv.py:10-35", "test/level2/symv/exo_symv_wrapper.h:22-32"),
    Proc("syr_rm", "syr",
         [enum("Uplo", UPLO_VALUES), size("n"), scalar("alpha"), win("x", "n"), win("x_copy", "n"),
          win("A", "n", "n", const=False)],
         "src/level2/syr.py:10-19", "test/level2/syr/exo_syr_wrapper.h:17-30"),
    Proc("syr2_rm", "syr2",
         [enum("Uplo", UPLO_VALUES), size("n"), scalar("alpha"), win("x", "n"), win("x_copy", "n"),
          win("y", "n"), win("y_copy", "n"), win("A", "n", "n", const=False)],
         "src/level2/syr2.py:9-27", "test/level2/syr2/exo_syr2_wrapper.h:20-34"),
    Proc("trmv_rm", "trmv",
         [enum("Uplo", UPLO_VALUES), enum("TransA", TRANS_VALUES), enum("Diag", DIAG_VALUES), size("n"),
          win("x", "n", const=False), win("A", "n", "n")],
         "src/level2/trmv.py:9-43", "test/level2/trmv/exo_trmv_wrapper.h:19-27"),
    Proc("trsv_rm", "trsv",
         [enum("Uplo", UPLO_VALUES), enum("TransA", TRANS_VALUES), enum("Diag", DIAG_VALUES), size("n"),
          win("x", "n", const=False), win("A", "n", "n")],
         "src/level2/trsv.py:10-52", "test/level2/trsv/exo_trsv_wrapper.h:19-27"),
    # ------------------------------------------------------------ banded level 2
    Proc("gbmv_row_major_NonTrans", "gbmv",
         [size("m"), size("n"), size("kl"), size("ku"), scalar("alpha"), scalar("beta"),
          win("a", "m", "kl + ku + 1"), win("x", "n"), win("y", "m", const=False)],
         "src/level2/gbmv.py:13-52", "test/level2/sgbmv/exo_sgbmv.h:9-19"),
    Proc("sbmv_scal_y", "sbmv", [size("n"), scalar("beta"), win("y", "n", const=False)],
         "src/level2/sbmv.py:12-16", "test/level2/ssbmv/exo_ssbmv.h:22-27"),
    Proc("sbmv_row_major_Upper", "sbmv",
         [size("n"), size("k"), scalar("alpha"), win("A", "n", "k + 1"), win("x", "n"),
          win("y", "n", const=False)],
         "src/level2/sbmv.py:19-33", "test/level2/ssbmv/exo_ssbmv.h:31-39"),
    Proc("sbmv_row_major_Lower", "sbmv",
         [size("n"), size("k"), scalar("alpha"), win("A", "n", "k + 1"), win("x", "n"),
          win("y", "n", const=False)],
         "src/level2/sbmv.py:36-50", "test/level2/ssbmv/exo_ssbmv.h:43-51"),
] + [
    Proc(f"tbmv_row_major_{uplo}_{tr}", "tbmv",
         [size("n"), size("k"), win("x", "n", const=False), win("A", "n", "k + 1"), size("Diag")],
         f"src/level2/tbmv.py:{lines}", f"test/level2/stbmv/exo_stbmv.h:{callers}")
    for uplo, tr, lines, callers in (("Upper", "NonTrans", "12-43", "20-28"), ("Lower", "NonTrans", "46-77", "46-54"),
                                     ("Upper", "Trans", "80-108", "32-40"), ("Lower", "Trans", "111-141", "58-66"))
] + [
    # the spec declares A as [n, n]; the band occupies n rows of k + 1 columns
    Proc(f"tbsv_row_major_{uplo}_{tr}", "tbsv",
         [size("n"), size("k"), win("x", "n", const=False), win("A", "n", "k + 1"), size("Diag")],
         f"src/level2/tbsv.py:{lines}", f"test/level2/stbsv/exo_stbsv.h:{callers}",
         variants=variants)
    for uplo, tr, lines, callers, variants in (
        ("Upper", "NonTrans", "12-45", "20-28", ("stride_any", "stride_1")),
        ("Lower", "NonTrans", "48-81", "39-47", ("stride_any", "stride_1")),
        ("Upper", "Trans", "84-105", "31-34", ("stride_any",)),
        ("Lower", "Trans", "108-129", "50-53", ("stride_any",)))
] + [
    # ------------------------------------------------------------ level 3
    Proc("gemm", "gemm",
         [enum("TransA", TRANS_VALUES), enum("TransB", TRANS_VALUES), size("M"), size("N"), size("K"),
          scalar("alpha"),
          win("A", "M", "K", used_when="TransA == 111"),
          win("B", "K", "N", used_when="TransB == 111"),
          win("AT", "K", "M", used_when="TransA != 111"),
          win("BT", "N", "K", used_when="TransB != 111"),
          win("C", "M", "N", const=False)],
         "src/level3/gemm.py:17-47", "test/level3/gemm/exo_gemm_wrapper.h:21-27"),
    Proc("mscal_rm", "mscal", [size("M"), size("N"), scalar("alpha"), win("A", "M", "N", const=False)],
         "src/level3/mscal.py:12-17", "test/level3/gemm/exo_gemm_wrapper.h:18-20"),
    Proc("trmscal_rm", "mscal",
         [enum("Uplo", UPLO_VALUES), size("N"), scalar("alpha"), win("A", "N", "N", const=False)],
         "src/level3/mscal.py:20-27", "test/level3/syrk/exo_syrk_wrapper.h:18-20"),
    Proc("symm_rm", "symm",
         [enum("Side", SIDE_VALUES), enum("Uplo", UPLO_VALUES), size("M"), size("N"), scalar("alpha"),
          win("Aleft", "M", "M", used_when="Side == 141"),
          win("Aright", "N", "N", used_when="Side != 141"),
          win("B", "M", "N"), win("C", "M", "N", const=False)],
         "src/level3/symm.py:16-47", "test/level3/symm/exo_symm_wrapper.h:21-26"),
    Proc("syrk_rm", "syrk",
         [enum("Uplo", UPLO_VALUES), enum("Trans", TRANS_VALUES), size("N"), size("K"), scalar("alpha"),
          win("A", "N", "K", used_when="Trans == 111"),
          win("A_", "N", "K", used_when="Trans == 111"),
          win("AT", "K", "N", used_when="Trans != 111"),
          win("AT_", "K", "N", used_when="Trans != 111"),
          win("C", "N", "N", const=False)],
         "src/level3/syrk.py:17-32", "test/level3/syrk/exo_syrk_wrapper.h:21-27"),
    Proc("syr2k_rm", "syr2k",
         [enum("Uplo", UPLO_VALUES), enum("Trans", TRANS_VALUES), size("N"), size("K"), scalar("alpha"),
          win("A", "N", "K", used_when="Trans == 111"),
          win("A_", "N", "K", used_when="Trans == 111"),
          win("B", "N", "K", used_when="Trans == 111"),
          win("B_", "N", "K", used_when="Trans == 111"),
          win("AT", "K", "N", used_when="Trans != 111"),
          win("AT_", "K", "N", used_when="Trans != 111"),
          win("BT", "K", "N", used_when="Trans != 111"),
          win("BT_", "K", "N", used_when="Trans != 111"),
          win("C", "N", "N", const=False)],
         "src/level3/syr2k.py:14-39", "test/level3/syr2k/exo_syr2k_wrapper.h:21-31"),
] + [
    Proc(f"trmm_row_major_Left_{uplo}_{tr}", "trmm",
         [size("m"), size("n"), scalar("alpha"), win("A", "m", "m"), win("B", "m", "n", const=False),
          size("Diag")],
         f"src/level3/trmm.py:{lines}", f"test/level3/strmm/exo_strmm.h:{callers}",
         precisions=("s",), variants=("",))
    for uplo, tr, lines, callers in (("Upper", "NonTrans", "12-23", "18-22"), ("Lower", "NonTrans", "26-46", "32-36"),
                                     ("Upper", "Trans", "49-69", "24-28"), ("Lower", "Trans", "72-83", "38-42"))
]

ORDER_VALUES = (101, 102)

# Native entry points with no counterpart in the reference: the split-
# precision and mixed-precision forms, and CBLAS-shaped gemm / gemv (storage
# order, beta fused, C / y not read when beta == 0).  Exported as exoblas_<p><name> (host or
# device pointers, synchronous) and exoblas_<p><name>_async (device pointers,
# stream-ordered), declared in include/exoblas_amd_ext.h.
EXT_PROCS: List[Proc] = [
    Proc("gemm_split", "ext",
         [enum("TransA", TRANS_VALUES), enum("TransB", TRANS_VALUES), size("M"), size("N"), size("K"),
          scalar("alpha"),
          win("A", "M", "K", used_when="TransA == 111"),
          win("B", "K", "N", used_when="TransB == 111"),
          win("AT", "K", "M", used_when="TransA != 111"),
          win("BT", "N", "K", used_when="TransB != 111"),
          win("C", "M", "N", const=False), enum("terms", (3, 4, 6))],
         "C += alpha op(A) op(B) with FP32 operands split into BF16 terms on the BF16 MFMA; "
         "arguments as exo_sgemm_stride_1 (src/level3/gemm.py:17-47) plus the number of BF16 products",
         "exoblas_amd.precision", precisions=("s",)),
    Proc("syrk_split", "ext",
         [enum("Uplo", UPLO_VALUES), enum("Trans", TRANS_VALUES), size("N"), size("K"), scalar("alpha"),
          win("A", "N", "K", used_when="Trans == 111"),
          win("AT", "K", "N", used_when="Trans != 111"),
          win("C", "N", "N", const=False), enum("terms", (3, 4, 6))],
         "one triangle of C += alpha A A^T (or A^T A) in the split BF16 form; "
         "arguments as exo_ssyrk_rm_stride_1 (src/level3/syrk.py:17-32) without the duplicate windows",
         "exoblas_amd.precision", precisions=("s",)),
    Proc("gemm_ex", "ext",
         [enum("Order", ORDER_VALUES), enum("TransA", TRANS_VALUES), enum("TransB", TRANS_VALUES),
          size("M"), size("N"), size("K"), scalar("alpha"),
          win("A", "((Order == 101) == (TransA == 111) ? M : K)", "((Order == 101) == (TransA == 111) ? K : M)"),
          win("B", "((Order == 101) == (TransB == 111) ? K : N)", "((Order == 101) == (TransB == 111) ? N : K)"),
          scalar("beta"),
          win("C", "(Order == 101 ? M : N)", "(Order == 101 ? N : M)", const=False)],
         "C = alpha op(A) op(B) + beta C with CBLAS dimensions in row-major (Order 101) or column-major (102) "
         "storage, C not read when beta == 0, beta applied in the product's epilogue; each window is the stored "
         "array (rows x leading dimension); the product as exo_<p>gemm_stride_1 (src/level3/gemm.py:17-47)",
         "exoblas_amd.cblas"),
    Proc("gemv_ex", "ext",
         [enum("Order", ORDER_VALUES), enum("TransA", TRANS_VALUES), size("M"), size("N"), scalar("alpha"),
          win("A", "(Order == 101 ? M : N)", "(Order == 101 ? N : M)"),
          win("x", "(TransA == 111 ? N : M)"), scalar("beta"),
          win("y", "(TransA == 111 ? M : N)", const=False)],
         "y = alpha op(A) x + beta y for an M x N matrix A in row-major (Order 101) or column-major (102) "
         "storage, y not read when beta == 0; element i of x at x.data + i * x.strides[0]; the product as "
         "exo_<p>gemv_rm_stride_any (src/level2/gemv.py:10-22)",
         "exoblas_amd.cblas"),
    Proc("trsv_mixed", "ext",
         [enum("Uplo", UPLO_VALUES), enum("TransA", TRANS_VALUES), enum("Diag", DIAG_VALUES), size("n"),
          win("x", "n", const=False), win("A", "n", "n"), enum("update", (16, 32)), size("refine")],
         "op(A) x = b with BF16 (update 16) or FP32 (update 32) off-diagonal updates, FP64 diagonal-block "
         "solves and `refine` steps of iterative refinement on an FP64 residual; arguments as "
         "exo_strsv_rm_stride_any (src/level2/trsv.py:10-52) plus update and refine",
         "exoblas_amd.precision", precisions=("s",)),
]

LIB_ORDER = ["asum", "axpy", "dot", "dsdot", "copy", "swap", "rot", "rotm", "scal",
             "gemv", "ger", "symv", "syr", "syr2", "trmv", "trsv", "gbmv", "sbmv", "tbmv", "tbsv",
             "gemm", "mscal", "symm", "syrk", "syr2k", "trmm"]


# ---------------------------------------------------------------- helpers

def elem_type(p: str, a: Arg) -> str:
    return a.ctype or PRECISIONS[p][0]


def win_struct(p: str, a: Arg) -> str:
    t = elem_type(p, a)
    tag = "f32" if t == "float" else "f64"
    return f"struct exo_win_{len(a.dims)}{tag}{'c' if a.const else ''}"


def c_param(p: str, a: Arg) -> str:
    if a.kind in ("size", "enum"):
        return f"{W} {a.name}"
    if a.kind == "scalar":
        return f"const {elem_type(p, a)} *{a.name}"
    if a.kind == "result":
        return f"{elem_type(p, a)} *{a.name}"
    if a.kind == "array":
        return f"const {elem_type(p, a)} *{a.name}"
    if a.kind == "win":
        return f"{win_struct(p, a)} {a.name}"
    raise ValueError(a.kind)


def symbol(proc: Proc, p: str, variant: Optional[str]) -> str:
    prefix = p if proc.fixed_prefix is None else proc.fixed_prefix
    if variant is None:
        return f"exoblas_{prefix}{proc.name}"
    if variant == "":
        return f"exo_{prefix}{proc.name}"
    return f"exo_{prefix}{proc.name}_{variant}"


def async_symbol(proc: Proc, p: str, variant: Optional[str]) -> str:
    prefix = p if proc.fixed_prefix is None else proc.fixed_prefix
    if not variant:
        return f"exoblas_{prefix}{proc.name}_async"
    return f"exoblas_{prefix}{proc.name}_{variant}_async"


def op_name(proc: Proc, p: str) -> str:
    if proc.fixed_prefix is not None or proc.lib == "ext":
        return f"exoblas::ops::{proc.name}"
    return f"exoblas::ops::{proc.name}<{PRECISIONS[p][0]}>"


def variants():
    for proc in PROCS:
        for p in proc.precisions:
            for v in proc.variants:
                yield proc, p, v


def ext_variants():
    for proc in EXT_PROCS:
        for p in proc.precisions:
            yield proc, p, None


def all_symbols() -> List[str]:
    return [symbol(proc, p, v) for proc, p, v in variants()]


def all_async_symbols() -> List[str]:
    return [async_symbol(proc, p, v) for proc, p, v in variants()]


# ---------------------------------------------------------------- emitters

HEADER_BANNER = "/* Generated by exoblas_amd/generator.py -- do not edit. */\n"


def emit_lib_header(lib: str) -> str:
    procs = [pr for pr in PROCS if pr.lib == lib]
    guard = f"EXOBLAS_AMD_EXO_{lib.upper()}_H"
    out = [HEADER_BANNER,
           f"/* Drop-in replacement for the Exo-generated exo_{lib}.h of the reference\n"
           f" * (library target EXO_BLAS::exo_{lib}).  Every function takes host or device\n"
           f" * pointers and returns once its results are visible to the caller.  */\n",
           f"#ifndef {guard}\n#define {guard}\n\n#include \"exo_types.h\"\n\n",
           "#ifdef __cplusplus\nextern \"C\" {\n#endif\n\n"]
    for proc in procs:
        out.append(f"/* {proc.name}: spec {proc.ref}; called from {proc.callers} */\n")
        for p in proc.precisions:
            for v in proc.variants:
                params = ", ".join(["void *ctxt"] + [c_param(p, a) for a in proc.args])
                out.append(f"void {symbol(proc, p, v)}({params});\n")
        out.append("\n")
    out.append("#ifdef __cplusplus\n}\n#endif\n\n#endif\n")
    return "".join(out)


def emit_ext_header() -> str:
    out = [HEADER_BANNER,
           "/* Native entry points of exoblas_amd with no counterpart in the reference\n"
           " * (the reference's exo_* functions are plain-precision only).  The plain\n"
           " * form takes host or device pointers and returns once results are\n"
           " * visible; the _async form takes device pointers and enqueues on\n"
           " * `stream`.  Both return 0 or an EXOBLAS_STATUS_* code (exoblas_amd.h). */\n",
           "#ifndef EXOBLAS_AMD_EXT_H\n#define EXOBLAS_AMD_EXT_H\n\n",
           "#include <hip/hip_runtime_api.h>\n#include \"exo_types.h\"\n\n",
           "#ifdef __cplusplus\nextern \"C\" {\n#endif\n\n"]
    for proc, p, v in ext_variants():
        out.append(f"/* {proc.ref} */\n")
        params = ", ".join(["void *ctxt"] + [c_param(p, a) for a in proc.args])
        out.append(f"int {symbol(proc, p, v)}({params});\n")
        params = ", ".join(["hipStream_t stream"] + [c_param(p, a) for a in proc.args])
        out.append(f"int {async_symbol(proc, p, v)}({params});\n\n")
    out.append("#ifdef __cplusplus\n}\n#endif\n\n#endif\n")
    return "".join(out)


def emit_async_header() -> str:
    out = [HEADER_BANNER,
           "/* Device-pointer, stream-ordered forms of every exo_* entry point.\n"
           " * Same arguments as the exo_* function of the same name, with the\n"
           " * hipStream_t the work is enqueued on in place of `void *ctxt`.  All window\n"
           " * data and result pointers must be device (or managed) memory; scalar\n"
           " * inputs are read on the host before the call returns.  Returns 0 on\n"
           " * success, otherwise an EXOBLAS_STATUS_* code (exoblas_amd.h).\n"
           " * Values that workgroups exchange through memory (reductions, split\n"
           " * gemv, trsv) carry a marker drawn from a per-stream counter on the\n"
           " * device, so a call captured in a hipGraph replays correctly.  A call\n"
           " * that needs more scratch than its stream holds allocates a larger\n"
           " * buffer (without synchronising; the old one stays valid for work\n"
           " * already queued or captured); during capture such growth is refused\n"
           " * with EXOBLAS_STATUS_NOT_SUPPORTED, so make the largest call once on\n"
           " * the stream before capturing.  Calls on one stream, and replays of\n"
           " * graphs captured on it, share its counters: serialise them.  Waits\n"
           " * between workgroups are bounded; exoblas_stream_check reports one\n"
           " * that expired (exoblas_amd.h). */\n",
           "#ifndef EXOBLAS_AMD_ASYNC_H\n#define EXOBLAS_AMD_ASYNC_H\n\n",
           "#include <hip/hip_runtime_api.h>\n#include \"exo_types.h\"\n\n",
           "#ifdef __cplusplus\nextern \"C\" {\n#endif\n\n"]
    for proc, p, v in variants():
        params = ", ".join(["hipStream_t stream"] + [c_param(p, a) for a in proc.args])
        out.append(f"int {async_symbol(proc, p, v)}({params});\n")
    out.append("\n#ifdef __cplusplus\n}\n#endif\n\n#endif\n")
    return "".join(out)


def _span_expr(a: Arg) -> Tuple[str, str]:
    """Element offset range [lo, hi) of a window, relative to .data."""
    if len(a.dims) == 1:
        return (f"exoblas::span_lo({a.dims[0]}, {a.name}.strides[0])",
                f"exoblas::span_hi({a.dims[0]}, {a.name}.strides[0])")
    return ("0", f"exoblas::span2_hi({a.dims[0]}, {a.dims[1]}, {a.name}.strides[0])")


def emit_body(proc: Proc, p: str, v: str, is_async: bool) -> str:
    sym = async_symbol(proc, p, v) if is_async else symbol(proc, p, v)
    lines = []
    lead = "hipStream_t stream" if is_async else "void *ctxt"
    params = ", ".join([lead] + [c_param(p, a) for a in proc.args])
    returns_int = is_async or v is None
    ret = "int" if returns_int else "void"
    lines.append(f"extern \"C\" {ret} {sym}({params}) {{")
    mode = "exoblas::Call::kAsync" if is_async else "exoblas::Call::kSync"
    src = "stream" if is_async else "ctxt"
    lines.append(f"  exoblas::Call call(\"{sym}\", {mode}, {src});")

    def R(expr):
        return f"return {expr};" if returns_int else f"return (void){expr};"
    # validation
    for a in proc.args:
        q = '"' + a.name + '"'
        if a.kind == "enum":
            vals = " || ".join(f"{a.name} == {x}" for x in a.values)
            lines.append(f"  if (!({vals})) " + R(f"call.fail_enum({q}, {a.name})"))
        if a.kind == "size":
            lines.append(f"  if ({a.name} < 0) " + R(f"call.fail_size({q}, {a.name})"))
    for a in proc.args:
        q = '"' + a.name + '"'
        if a.kind == "win" and len(a.dims) == 2:
            use = a.used_when or "true"
            lines.append(f"  if (({use}) && {a.name}.strides[1] != 1) "
                         + R(f"call.fail_stride({q}, {a.name}.strides[1])"))
            lines.append(f"  if (({use}) && {a.dims[0]} > 1 && {a.name}.strides[0] < {a.dims[1]}) "
                         + R(f"call.fail_ld({q}, {a.name}.strides[0])"))
    # windows
    idx = 0
    win_index = {}
    for a in proc.args:
        if a.kind == "win":
            lo, hi = _span_expr(a)
            use = a.used_when or "true"
            t = elem_type(p, a)
            lines.append(f"  call.window({idx}, {a.name}.data, {lo}, {hi}, sizeof({t}), "
                         f"{'false' if a.const else 'true'}, {use});")
            win_index[a.name] = idx
            idx += 1
        elif a.kind == "result":
            t = elem_type(p, a)
            lines.append(f"  call.window({idx}, {a.name}, 0, 1, sizeof({t}), true, true, true);")
            win_index[a.name] = idx
            idx += 1
    lines.append(f"  if (!call.stage()) {R('call.status()')}")
    # scalars / arrays
    for a in proc.args:
        t = elem_type(p, a)
        if a.kind == "scalar":
            lines.append(f"  const {t} {a.name}_v = call.scalar({a.name});")
        if a.kind == "array":
            lines.append(f"  {t} {a.name}_v[{a.length}];")
            lines.append(f"  call.array({a.name}, {a.name}_v, {a.length});")
    lines.append(f"  if (!call.ok()) {R('call.status()')}")
    # op call
    cargs = ["call.ctx()"]
    for a in proc.args:
        t = elem_type(p, a)
        if a.kind in ("size", "enum"):
            cargs.append(f"(int64_t){a.name}")
        elif a.kind == "scalar":
            cargs.append(f"{a.name}_v")
        elif a.kind == "array":
            cargs.append(f"{a.name}_v")
        elif a.kind == "result":
            cargs.append(f"call.dev<{t}>({win_index[a.name]})")
        elif a.kind == "win":
            cq = "const " if a.const else ""
            cargs.append(f"call.dev<{cq}{t}>({win_index[a.name]})")
            for k in range(len(a.dims)):
                if k == 0 or len(a.dims) == 1:
                    cargs.append(f"(int64_t){a.name}.strides[{k}]")
    lines.append(f"  call.run([&] {{ {op_name(proc, p)}({', '.join(cargs)}); }});")
    lines.append(f"  {R('call.finish()')}")
    lines.append("}")
    return "\n".join(lines) + "\n"


def emit_exports() -> str:
    out = ["// Generated by exoblas_amd/generator.py -- do not edit.\n",
           "// extern \"C\" definitions of every exo_* symbol and its _async form.\n",
           "#include \"runtime.hpp\"\n#include \"ops.hpp\"\n"]
    for lib in LIB_ORDER:
        out.append(f"#include \"exo_{lib}.h\"\n")
    out.append("#include \"exoblas_amd_async.h\"\n#include \"exoblas_amd_ext.h\"\n\n")
    for proc, p, v in list(variants()) + list(ext_variants()):
        out.append(f"// {proc.ref}\n")
        out.append(emit_body(proc, p, v, False))
        out.append(emit_body(proc, p, v, True))
        out.append("\n")
    return "".join(out)


def emit_symbol_list() -> str:
    ext = [f(proc, p, v) for proc, p, v in ext_variants() for f in (symbol, async_symbol)]
    return "\n".join(all_symbols() + all_async_symbols() + ext) + "\n"


def outputs():
    files = {}
    for lib in LIB_ORDER:
        files[os.path.join(ROOT, "include", f"exo_{lib}.h")] = emit_lib_header(lib)
    files[os.path.join(ROOT, "include", "exoblas_amd_async.h")] = emit_async_header()
    files[os.path.join(ROOT, "include", "exoblas_amd_ext.h")] = emit_ext_header()
    files[os.path.join(ROOT, "exoblas_amd", "csrc", "gen_exports.cpp")] = emit_exports()
    files[os.path.join(ROOT, "exoblas_amd", "csrc", "symbols.txt")] = emit_symbol_list()
    return files


def write_all() -> List[str]:
    changed = []
    for path, text in outputs().items():
        old = None
        if os.path.exists(path):
            with open(path) as f:
                old = f.read()
        if old != text:
            with open(path, "w") as f:
                f.write(text)
            changed.append(path)
    return changed


if __name__ == "__main__":
    ch = write_all()
    print(f"generator: {len(ch)} file(s) rewritten", file=sys.stderr)
