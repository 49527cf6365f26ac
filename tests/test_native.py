"""CPU builds of host code with AddressSanitizer + UndefinedBehaviorSanitizer.

tests/native/Makefile compiles the library's host-side input handling
(gf2_amd/csrc/host_codes.cpp: the from_coo CSR build of gf2l_code_edges, the
GF2DATA reader of gf2l_code_gf2data, the DVB table checks) and the CPU oracle
with -fsanitize=address,undefined.  The drivers feed hostile sizes (rows or
cols = 2^32 - 1 in GF2DATA headers and edge lists, the out-of-bounds write of
an n + 1 that wrapped to 0), truncated and randomly mutated files, and expect
error codes; any sanitizer report aborts the program and fails the test.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")


@pytest.fixture(scope="module")
def sanitize_build():
    subprocess.run(["make", "-s", "-C", NATIVE, "sanitize"], check=True)
    return os.path.join(NATIVE, "_bin")


def _run(path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([path], capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr
    return p.stdout


def test_host_input_handling_under_sanitizers(sanitize_build):
    assert "all expectations hold" in _run(os.path.join(sanitize_build, "host_sanitize"))


def test_oracle_under_sanitizers(sanitize_build):
    assert "all expectations hold" in _run(os.path.join(sanitize_build, "oracle_sanitize"))
