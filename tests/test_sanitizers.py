"""AddressSanitizer and UndefinedBehaviorSanitizer over the host code, on the
CPU: the oracle (oracle/liboracle_san.so, built by `make -C oracle
liboracle_san.so`) under its own test suite, and the library's host code
(the planner, argument checks, validity packing, pandas conventions) built
by `make -C cudf_amd/csrc sanitize` under the CPU test suite. Each suite runs
in a child process with the sanitizer runtime preloaded; any report ends the
child with a failure."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gcc_runtime(name: str) -> str | None:
    try:
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _run_child(env_extra: dict, args: list) -> subprocess.CompletedProcess:
    env = dict(os.environ)
    env.update(env_extra)
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] + args,
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)


def test_oracle_under_asan_and_ubsan():
    asan, ubsan = _gcc_runtime("libasan.so"), _gcc_runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "liboracle_san.so"], check=True)
    lib = os.path.join(ROOT, "oracle", "liboracle_san.so")
    r = _run_child({"LD_PRELOAD": f"{asan}:{ubsan}", "MGB_ORACLE_LIB": lib},
                   ["tests/test_oracle.py", "tests/test_oracle_collect.py",
                    "tests/test_sanitizers.py::test_sanitized_oracle_is_loaded"])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def test_sanitized_oracle_is_loaded():
    """Run by the child of test_oracle_under_asan_and_ubsan: the oracle it
    tests is the sanitizer build."""
    if not os.environ.get("MGB_ORACLE_LIB"):
        pytest.skip("only meaningful inside the sanitizer run")
    from oracle import oracle as O

    O.load()
    maps = open("/proc/self/maps").read()
    assert "liboracle_san.so" in maps and "libasan" in maps


def test_library_host_code_under_asan_and_ubsan():
    lib = os.path.join(ROOT, "cudf_amd", "lib", "san", "libcudf_amd.so")
    rts = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rts or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc or clang's sanitizer runtime not installed")
    # incremental: rebuilds only what changed since the last sanitizer build
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "cudf_amd", "csrc"), "sanitize"], check=True,
                   capture_output=True, timeout=1800)
    r = _run_child({"LD_PRELOAD": rts[-1], "MGB_LIB_PATH": lib},
                   ["tests/test_capi.py", "tests/test_pandas_accel.py",
                    "tests/test_sanitizers.py::test_sanitized_library_is_loaded"])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def test_sanitized_library_is_loaded():
    """Run by the child of test_library_host_code_under_asan_and_ubsan."""
    if not os.environ.get("MGB_LIB_PATH"):
        pytest.skip("only meaningful inside the sanitizer run")
    from cudf_amd import _lib

    _lib.load()
    maps = open("/proc/self/maps").read()
    assert "lib/san/libcudf_amd.so" in maps and "asan" in maps
