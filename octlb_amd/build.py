"""Build liboctlb_gpu.so in-tree with hipcc for gfx950.

    python -m octlb_amd.build

FP64 parity needs the reference's rounding: no FMA contraction
(-ffp-contract=off) and no fast-math.  Packed FP32 instructions are turned
off for every HIP translation unit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "liboctlb_gpu.so"
ARCH = os.environ.get("OCTLB_OFFLOAD_ARCH", "gfx950")

SOURCES = ["lbm.hip", "api.hip", "links.hip", "adjoint.hip", "migrate.hip", "voxelize.hip", "forest.cpp",
           "vtk_out.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_ONLY = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        cmd = [cc] + COMMON
        if src.endswith(".hip"):
            cmd += HIP_ONLY + ["-x", "hip"]
        cmd += ["-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


def build_oracle(verbose: bool = False) -> None:
    """Compile the test-only checkers under oracle/ (the restatement and,
    where the reference sources exist, the reference harness)."""
    cmd = ["make", "-C", str(ROOT / "oracle"), "-j4", "all"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    v = "-v" in sys.argv
    print(build(verbose=v))
    if "--no-oracle" not in sys.argv:
        build_oracle(verbose=v)
