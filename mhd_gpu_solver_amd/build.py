"""Build recipe for libmhd_hip.so (hipcc, gfx950). Run: python -m mhd_gpu_solver_amd.build

Compiles every translation unit under csrc/ and links the shared library in-tree next to this
file, so that it travels with the source tree and is the object the package loads.

* mhd_fp64.hip, mhd_solver.hip, mhd_host.cpp, ... : -ffp-contract=off (fp64 parity with the
  CPU path depends on every product and sum rounding separately)
* mhd_fp32.hip (single-run FP32 kernels) and mhd_fp64fast.hip (the fast fp64 stepping
  kernels): contraction within each source expression (-ffp-contract=on), plus the fused
  multiply-adds the sources spell out. Not -ffp-contract=fast: across expressions the compiler
  picks which product of a sum to fuse from the code around it, which differs between the two
  directions these kernels march row segments in, and the results must not depend on the
  direction (nor, through it, on the tiling).
* mhd_fp32batch.hip (the batch engine's FP32 kernels, which march in one direction):
  contraction across expressions (-ffp-contract=fast).
* all HIP code: packed fp32 instructions disabled (-packed-fp32-ops)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
INCLUDE = ROOT / "include"
LIB = PKG / "libmhd_hip.so"
OBJDIR = PKG / "_build"
ARCH = os.environ.get("MHD_HIP_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CONTRACT_ON = {"mhd_fp32.hip", "mhd_fp64fast.hip"}
CONTRACT_FAST = {"mhd_fp32batch.hip"}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(path).exists():
        raise RuntimeError("hipcc not found: the hip backend needs ROCm's hipcc to build")
    return path


def sources():
    return sorted(p for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp"))


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = OBJDIR / (src.name + ".o")
    contract = ("-ffp-contract=fast" if src.name in CONTRACT_FAST else
                "-ffp-contract=on" if src.name in CONTRACT_ON else "-ffp-contract=off")
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [hipcc(), *COMMON, contract, *lang, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> pathlib.Path:
    OBJDIR.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
