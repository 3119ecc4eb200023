"""The pass layout's index maps (pixel order of the tiles, queue slot ->
pixel and sample) are bijections: built for the host from the same header the
kernels use and checked over image sizes with ragged edges, 1080p included.
A map that is not one would make the film kernels write outside the film."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_pixel_order_and_slot_maps_are_bijections(tmp_path):
    exe = tmp_path / "order_check"
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", f"-I{ROOT / 'include'}", f"-I{ROOT / 'raycer_amd' / 'csrc'}",
           str(ROOT / "tests" / "native" / "order_check.hip"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("0 failures")
