"""The host half of the library and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (a host-only g++ build: python -m raycer_amd.build
--sanitize DIR).  The test program, tests/native/host_sanitize.cpp, loads the
reference's meshes and sample scene, round-trips scene and BVH files, feeds the
scene reader every truncation and seeded mutations of the sample and hostile
nesting, builds BVHs with both builders and runs the oracle's traversal,
renderer, split-tail estimator and MLP.  Any sanitizer finding aborts it."""
import os
import shutil
import subprocess

import pytest

from raycer_amd import build as rb
from tests.conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_host_code_and_oracle_are_clean_under_asan_and_ubsan(tmp_path):
    exe = rb.build_sanitized(tmp_path / "san")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(ROOT / "tests" / "golden"), str(tmp_path)], capture_output=True, text=True,
                       timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.stdout.strip().splitlines()[-1].endswith(", 0 failures")
