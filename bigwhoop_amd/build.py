"""Build libbwc.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

    python -m bigwhoop_amd.build

The library lands in bigwhoop_amd/lib/libbwc.so. Every HIP object is built
for gfx950 only, with -ffp-contract=off (no FMA contraction: the FP64 DWT must
reproduce the reference's rounding) and packed FP32 disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "lib"
OBJ = PKG / "_build"
LIB = OUT / "libbwc.so"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-I" + str(PKG.parent / "include")]
# -structurizecfg-skip-uniform-regions: branches whose condition is the same
# on every lane stay scalar branches instead of being rewritten into
# exec-masked flow blocks; the tier-1 decoder's control flow is uniform and
# halves in size with it
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-mllvm", "-structurizecfg-skip-uniform-regions=true"]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", str(src),
                                  "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    OUT.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs] + \
          ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
