"""The host fuzz tests again, against the library's host code built with
AddressSanitizer and UBSan (lensed_amd.build.build_sanitized): FITS reader
and writer, data preparation, INI/option parsing.  They run in a child
process with the sanitizer runtime preloaded; any report fails the test."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT
from lensed_amd import build as B

FUZZ = ["tests/test_fits.py", "tests/test_config.py"]


@pytest.mark.skipif(B.asan_runtime() is None, reason="no shared ASan runtime in this toolchain")
def test_fuzzers_under_address_and_ub_sanitizers():
    if not os.path.exists(B.ASAN_LIB):
        B.build_sanitized()
    env = dict(os.environ, LENSED_LIB=B.ASAN_LIB, LD_PRELOAD=B.asan_runtime(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + FUZZ
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert "passed" in out
