"""Host code under the address and undefined-behaviour sanitizers (CPU only, no GPU).

* The library's host-side set-up (bb_host.hpp: the stage kernel's index and matrix-core
  tables of every degree, solver and batch set-up with every rejected input, the exception
  guard of the C entry points) is built host-only and run by tests/native/host_check.cpp.
* The oracle (oracle/bbref.c) is built with the sanitizers and the oracle's own test files
  run against that build, with the sanitizer runtime preloaded into the interpreter.
* A planted out-of-bounds write through the sanitized oracle must be reported, so a build
  that silently lost its instrumentation cannot pass.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bernsteinbasis.jl_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_oracle(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not present")
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("sanitizer runtimes not present")
    import oracle

    so = str(tmp_path_factory.mktemp("san") / "liboracle_san.so")
    oracle.build_sanitized(so)
    env = dict(os.environ)
    env.update({"BB_ORACLE_SO": so, "LD_PRELOAD": f"{asan}:{ubsan}",
                "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    return env


def test_host_setup_under_sanitizers(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not present")
    exe = str(tmp_path / "host_check")
    cmd = [HIPCC, "--offload-host-only", "-x", "hip", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, os.path.join(ROOT, "tests", "native", "host_check.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "host checks passed" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])


def test_oracle_tests_under_sanitizers(san_oracle):
    files = [os.path.join(ROOT, "tests", f) for f in
             ("test_oracle_operators.py", "test_oracle_rhs.py", "test_oracle_mixed_degree.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *files],
                       capture_output=True, text=True, timeout=900, env=san_oracle, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout


def test_sanitized_oracle_reports_a_planted_overflow(san_oracle):
    """orc_l_j writes 20 values; a 10-value buffer must be reported as a heap overflow"""
    code = ("import sys, numpy as np; sys.path.insert(0, %r); import oracle; "
            "oracle.lib().orc_l_j(5, np.zeros(10)); print('not caught')") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, env=san_oracle)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, (r.stdout[-1000:], r.stderr[-3000:])
