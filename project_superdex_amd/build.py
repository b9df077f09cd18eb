"""Build libsdxfem.so (HIP kernels + host driver + C ABI) for gfx950 in-tree.

Run as ``python -m project_superdex_amd.build`` or via __graft_entry__.build().
Every HIP translation unit is compiled with packed FP32 instructions turned
off (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsdxfem.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("SDX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["sdx_kernels.hip", "sdx_mlp.hip", "sdx_world.cpp", "sdx_api.cpp", "sdx_pcg.cpp", "sdx_mlp_host.cpp",
           "sdx_pcg_noalloc.hip", "sdx_capture.cpp"]
HEADERS = ["sdx_kernels.h", "sdx_layout.h", "sdx_maps.h", "sdx_math.h", "sdx_world.h", "sdx_device.h", "sdx_mlp.h", "sdx_mlp_host.h", "sdx_pcg.h", "sdx_capture.h", "sdx_rowblocks.h"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; ROCm is required to build libsdxfem.so")
    return path


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
