"""Every compile-time switch of the kernels still compiles in its non-default
setting.  The switches keep the measured alternatives of the traversal, the
queues and the launch shapes (node format and order, stack split, hit-test
records, child order, leaf postponing, parking rule, camera fusion and
packets, XCD remap); the library is built with the defaults only, so without
this check a switched-off path could stop compiling unnoticed.  hipcc runs in
syntax-only mode for gfx950 (templates are instantiated and checked, no code
is generated)."""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from raycer_amd import build

SWITCHES = [
    "RC_QNODE=0", "RC_LDS_TOP=1", "RC_TRI_EDGES=1", "RC_REG_STACK=0", "RC_REG_STACK=2", "RC_PIXEL_MAJOR=0",
    "RC_FAST_RCP=0", "RC_SLAB_SELECT=0", "RC_NT_QUEUES=1", "RC_TRAV_IFIF=0", "RC_ANY_SORT=1",
    "RC_LEAF_POSTPONE=0", "RC_FUSED_CAMERA=0", "RC_XCD_REMAP=1", "RC_PARK_KEEP=0", "RC_PACKET_CAMERA=1",
    "RC_NODE_ORDER=1",
]


def _syntax_check(define: str) -> tuple[str, int, str]:
    sources = sorted(str(p) for p in build.CSRC.glob("*.hip"))
    cmd = [build.HIPCC, "--offload-arch=gfx950", "-std=c++17", "-fsyntax-only", f"-I{build.ROOT / 'include'}",
           f"-I{build.CSRC}", f"-D{define}", *sources]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    errors = "\n".join(l for l in p.stdout.splitlines() if "error" in l)
    return define, p.returncode, errors


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc not installed")
def test_every_switch_compiles_in_its_other_setting():
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_syntax_check, SWITCHES))
    failed = [f"{d}: rc {rc}\n{err}" for d, rc, err in results if rc != 0]
    assert not failed, "\n".join(failed)


def test_the_switch_list_names_only_switches_the_sources_define():
    text = "".join(p.read_text() for p in build.CSRC.glob("*.hip")) + "".join(
        p.read_text() for p in build.CSRC.glob("*.hpp"))
    for s in SWITCHES:
        assert f"#ifndef {s.split('=')[0]}" in text, s
