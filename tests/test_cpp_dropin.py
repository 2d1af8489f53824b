"""The C++ drop-in (cudf::groupby::groupby over libcudf_amd.so): the test
program tests/cpp/test_dropin.cpp runs the reference's groupby fixtures
through the cudf C++ API."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_dropin")


def test_dropin_program_links_the_library():
    assert os.path.exists(BIN), "build it with __graft_entry__.build()"
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True, check=True).stdout
    line = [ln for ln in out.splitlines() if "libcudf_amd.so" in ln]
    assert line and "not found" not in line[0]


@pytest.mark.gpu
def test_dropin_program(gpu):
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failures" in p.stdout
