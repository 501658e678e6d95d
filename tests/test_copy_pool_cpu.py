"""The host copy pool of the pinned staging path (exoblas_amd/csrc/copy_pool.hpp),
built with g++ and run on the CPU: exact copies, no writes past the end,
concurrent callers."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "exoblas_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_copy_pool(tmp_path):
    exe = tmp_path / "copy_pool_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", CSRC,
                    os.path.join(HERE, "native", "copy_pool_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
