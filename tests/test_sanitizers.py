"""The host C++ under AddressSanitizer and UndefinedBehaviorSanitizer, on the
CPU: the library's tree topology and STL code driven by
baby_shark_amd/csrc/host_check.cpp, and the oracle (the CPU restatement the
GPU tests check against) running its own pinning tests.  A sanitizer report
ends the run with a non-zero status."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed for the sanitizer builds")


def _run(cmd, env=None, timeout=900):
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"{' '.join(cmd)} failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    return r


def test_host_code_under_sanitizers():
    _run(["make", "-s", "-C", "baby_shark_amd/csrc", "host_check"])
    r = _run([os.path.join("baby_shark_amd", "csrc", "_build", "host_check"), os.path.join("tests", "golden")],
             env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                      UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"))
    assert "host checks passed" in r.stdout


def test_oracle_under_sanitizers():
    _run(["make", "-s", "-C", "oracle", "asan"])
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["g++", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, BS_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_asan.so"),
               LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    tests = ["tests/test_oracle.py", "tests/test_oracle_tree.py", "tests/test_oracle_offset.py",
             "tests/test_oracle_dc.py", "tests/test_oracle_mc.py", "tests/test_oracle_bvh.py",
             "tests/test_oracle_builder.py", "tests/test_golden.py"]
    _run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + tests, env=env,
         timeout=1800)
