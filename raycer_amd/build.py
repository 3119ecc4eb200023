"""Builds libraycer.so (gfx950 HIP kernels + C ABI + host C++) in-tree.

    python -m raycer_amd.build            # library + console runner
    python -m raycer_amd.build --oracle   # also the FP64 oracle (test infrastructure)

Every HIP file is compiled by hipcc for gfx950 only, with packed-FP32
instruction selection disabled.  Objects go to raycer_amd/_build/, the
library to raycer_amd/libraycer.so and the C++ console runner to
raycer_amd/raycer_render.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libraycer.so"
RUNNER = PKG / "raycer_render"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def _objects(build_dir: Path, defines: list[str]) -> list[tuple[list[str], Path]]:
    jobs = []
    dflags = [f"-D{d}" for d in defines]
    for src in sorted(CSRC.glob("*.hip")):
        obj = build_dir / (src.stem + ".hip.o")
        jobs.append(([HIPCC, *HIP_FLAGS, *dflags, "-c", str(src), "-o", str(obj)], obj))
    for src in sorted((CSRC / "host").glob("*.cpp")):
        obj = build_dir / (src.stem + ".cpp.o")
        jobs.append((["g++", *CXX_FLAGS, *dflags, "-c", str(src), "-o", str(obj)], obj))
    return jobs


def build_library(jobs: int = 8, defines: list[str] | None = None, out_dir: Path | None = None) -> Path:
    """Builds libraycer.so; with defines (NAME=VALUE, the kernels' compile-time
    switches such as RC_REG_STACK) into out_dir, for A/B measurements (load
    it with RAYCER_LIBRARY=<out_dir>/libraycer.so)."""
    defines = list(defines or [])
    lib = (out_dir / "libraycer.so") if out_dir else LIB
    build_dir = (out_dir / "_build") if out_dir else BUILD
    build_dir.mkdir(parents=True, exist_ok=True)
    work = _objects(build_dir, defines)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda j: _run(j[0]), work))
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *[str(o) for _, o in work]])
    os.replace(tmp, lib)
    return lib


def build_runner() -> Path:
    src = CSRC / "cpp" / "console_runner.cpp"
    if not src.exists():
        return RUNNER
    _run(["g++", *CXX_FLAGS, str(src), "-o", str(RUNNER), f"-L{PKG}", "-lraycer", "-Wl,-rpath,$ORIGIN"])
    return RUNNER


SANITIZE_FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                  "-fno-omit-frame-pointer", "-fopenmp", "-pthread"]


def build_sanitized(out_dir: Path) -> Path:
    """The host half of the library (csrc/host: OBJ/PLY, scene files, BVH
    builders and files, scene generators) and the FP64 oracle, built with
    AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) and
    linked into the test program tests/native/host_sanitize.cpp."""
    out_dir.mkdir(parents=True, exist_ok=True)
    exe = out_dir / "host_sanitize"
    srcs = [ROOT / "tests" / "native" / "host_sanitize.cpp", *sorted((CSRC / "host").glob("*.cpp")),
            *(ROOT / "oracle" / f for f in ("bvh.cpp", "render.cpp", "mlp.cpp"))]
    _run(["g++", *SANITIZE_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", f"-I{CSRC / 'host'}", f"-I{ROOT / 'oracle'}",
          *[str(x) for x in srcs], "-o", str(exe)])
    return exe


def build_oracle() -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "-j4"])


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--define", action="append", default=[], help="NAME=VALUE for a variant build (with --out)")
    ap.add_argument("--out", type=Path, default=None, help="directory of a variant build's libraycer.so")
    ap.add_argument("--sanitize", type=Path, default=None,
                    help="build the host half and the oracle with ASan + UBSan into this directory (host_sanitize)")
    a = ap.parse_args(argv)
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        print("hipcc not found", file=sys.stderr)
        return 1
    if a.sanitize:
        print(f"built {build_sanitized(a.sanitize.resolve())}")
        return 0
    if a.define and not a.out:
        print("--define needs --out (the in-tree library is the default build)", file=sys.stderr)
        return 1
    if a.out:
        print(f"built {build_library(a.j, a.define, a.out.resolve())}")
        return 0
    build_library(a.j)
    build_runner()
    if a.oracle:
        build_oracle()
    print(f"built {LIB}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
