"""Build recipe for the in-tree native libraries.

    python -m segmentctscans_amd._build          # build everything

Outputs (git-ignored, shipped with the tree):
    segmentctscans_amd/_lib/libsct.so           HIP library (gfx950), C ABI in include/sct.h
    segmentctscans_amd/_lib/libsct_phantom.so   host phantom generator (OpenMP)
    oracle/_build/libsct_oracle.so              CPU oracle (test infrastructure)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "_lib"
OBJ = PKG / "_build"

HIP_SOURCES = ["context.hip", "potts.hip", "potts_fused.hip", "chambolle.hip", "reductions.hip", "selftest.hip", "gap.hip", "batch.hip", "adjoint.hip", "archive.hip", "gmm.hip", "bimm.hip"]

# -ffp-contract=off: the only fused multiply-adds are the explicit __fmaf_rn
# calls, so results match the FP32 host mirror bit for bit.  Packed FP32
# instructions are disabled (target feature -packed-fp32-ops).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout + proc.stderr)


def _compile_hip(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    _run([_hipcc(), *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(out)])
    return out


def build_hip() -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    srcs = [s for s in HIP_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    target = LIB / "libsct.so"
    tmp = LIB / "libsct.so.tmp"
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, target)
    return target


def build_phantom() -> Path:
    LIB.mkdir(exist_ok=True)
    target = LIB / "libsct_phantom.so"
    tmp = LIB / "libsct_phantom.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-o", str(tmp), str(CSRC / "phantom.cpp")])
    os.replace(tmp, target)
    return target


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "libsct_oracle.so"


def build_cpp_example() -> Path:
    """tests/cpp/api_example (device) and archive_example (host only): the C++
    API (include/segmentctscans.hpp) over libsct.so."""
    out_dir = ROOT / "tests" / "cpp" / "_build"
    out_dir.mkdir(exist_ok=True)
    for name in ("api_example", "archive_example"):
        _run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / f"{name}.cpp"),
              "-L", str(LIB), "-lsct", "-Wl,-rpath,$ORIGIN/../../../segmentctscans_amd/_lib", "-o",
              str(out_dir / name)])
    return out_dir / "api_example"


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_hip), ex.submit(build_phantom), ex.submit(build_oracle)]
        for f in futs:
            f.result()
    build_cpp_example()


if __name__ == "__main__":
    build_all()
    print("built", LIB / "libsct.so", LIB / "libsct_phantom.so", file=sys.stderr)
