"""Build recipe for the native library (libgs.so) of this package.

Compiles the gfx950 kernels, the C ABI and the host-side producer with
hipcc into ``_lib/libgs.so`` next to this file, and, where Node's N-API
headers are installed, the Node addon ``_lib/gs_napi.node`` over it.
Floating-point contraction is disabled everywhere (the kernels must round
exactly like the CPU restatement they are checked against), and packed FP32
instructions are disabled for the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB_DIR = HERE / "_lib"
LIB_PATH = LIB_DIR / "libgs.so"
ROOT = HERE.parent

SOURCES = ["gs_kernels.hip", "gs_cull.hip", "gs_api.hip", "gs_producer.cpp"]
HEADERS = ["gs_internal.h", "gs_project.h", "gs_math.h"]

ARCH = os.environ.get("GS_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Matrix-core results in ordinary VGPRs: the cull MLP's epilogues read
    # every accumulator, and AGPR results would cost a copy each.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgs")


NODE_INCLUDE = Path(os.environ.get("GS_NODE_INCLUDE", "/usr/include/node"))
NAPI_PATH = LIB_DIR / "gs_napi.node"


def build_napi(verbose: bool = False) -> Path | None:
    """Compile node/gs_napi.cpp against libgs.so; None when N-API is absent."""
    if not (NODE_INCLUDE / "node_api.h").exists():
        return None
    tmp = LIB_DIR / "gs_napi.node.tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{NODE_INCLUDE}",
           str(HERE / "node" / "gs_napi.cpp"), "-o", str(tmp), f"-L{LIB_DIR}", "-lgs",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"g++ failed on the N-API addon ({proc.returncode}):\n{proc.stderr[-8000:]}")
    os.replace(tmp, NAPI_PATH)
    return NAPI_PATH


def build(verbose: bool = False) -> Path:
    """Compile every source into _lib/libgs.so (always recompiles)."""
    LIB_DIR.mkdir(exist_ok=True)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *[str(CSRC / s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr[-8000:]}")
    warnings = [l for l in proc.stderr.splitlines() if "warning" in l and "packed-fp32" not in l]
    if verbose and warnings:
        print("\n".join(warnings), file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    build_napi(verbose)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
