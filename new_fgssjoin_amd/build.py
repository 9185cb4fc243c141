"""Build the native library and CLI in-tree with hipcc for gfx950.

Outputs (git-ignored, shipped with the tree):
  new_fgssjoin_amd/libfgssjoin.so   C ABI declared in include/fgssjoin.h
  new_fgssjoin_amd/fgssjoin         command-line tool (links the library)

Every translation unit is compiled with -ffp-contract=off and without
fast-math: the float32 similarity bounds must round exactly like the
reference's.  Device code also disables packed fp32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfgssjoin.so")
CLI = os.path.join(HERE, "fgssjoin")
ARCH = os.environ.get("FGSS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build new_fgssjoin_amd")


COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-I", os.path.join(ROOT, "include")]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("collection.cpp", False),
    ("join.hip", True),
    ("tokenize.hip", True),
]


def _compile(src: str, device: bool, verbose: bool) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + (DEVICE if device else []) + ["-c", os.path.join(CSRC, src), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    """Compile everything (always; this is the build check)."""
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s[0], s[1], verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libfgssjoin.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cli_obj = _compile("cli.cpp", False, verbose)
    cmd = [hipcc(), "-o", CLI, cli_obj, "-L", HERE, "-lfgssjoin", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking fgssjoin CLI failed:\n{r.stdout}\n{r.stderr}")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
