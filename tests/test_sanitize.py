"""The host-only C++ of the library (deck reader, writers, geometry, error
state) and the oracle, built with AddressSanitizer and
UndefinedBehaviorSanitizer (the `sanitize` targets of
fasterhenry_amd/csrc/Makefile and oracle/Makefile) and driven through valid,
truncated, mutated and hostile inputs by tests/sanitize/*_driver.*.  CPU
only; no GPU code is built with a sanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")

ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _build_and_run(directory, binary, *args):
    subprocess.run(["make", "-C", os.path.join(ROOT, directory), "sanitize"], check=True,
                   capture_output=True, timeout=600)
    exe = os.path.join(ROOT, "_build", "sanitize", binary)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=ENV,
                       cwd=ROOT)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all checks passed" in r.stdout


def test_host_code_under_asan_and_ubsan():
    _build_and_run(os.path.join("fasterhenry_amd", "csrc"), "host_driver",
                   os.path.join(ROOT, "tests", "golden", "decks"))


def test_oracle_under_asan_and_ubsan():
    _build_and_run("oracle", "oracle_driver")
