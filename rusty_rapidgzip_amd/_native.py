"""Build and load the native libraries of this package.

``librrg_gpu.so``   HIP kernels + host driver + C ABI (include/rrg_gpu.h), built by hipcc for gfx950.
``librrg_synth.so`` seeded workload generators and gzip/BGZF writers (gcc + zlib).

Libraries are built in-tree so that they travel with the source tree.  Loading
fails loudly when a library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
GPU_LIB = os.path.join(PKG_DIR, "librrg_gpu.so")
SYNTH_LIB = os.path.join(PKG_DIR, "synth", "librrg_synth.so")
CLI_BIN = os.path.join(PKG_DIR, "rapidgzip-gpu")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are disabled for all device code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd)}):\n{r.stdout}\n{r.stderr}")


def build_gpu_lib() -> str:
    srcs = [os.path.join(CSRC, f) for f in ("rrg_kernels.hip", "rrg_gpu.cpp", "rrg_serve.cpp", "rrg_slab.cpp",
                                              "rrg_noalloc.hip")]
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(REPO_DIR, "include"), *srcs, "-o", tmp, "-lpthread"])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_cli() -> str:
    """The command-line drop-in (cli/rrg_cli.cpp), linked against librrg_gpu.so next to it."""
    src = os.path.join(PKG_DIR, "cli", "rrg_cli.cpp")
    tmp = CLI_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(REPO_DIR, "include"), src, "-o", tmp, "-L", PKG_DIR,
          "-lrrg_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, CLI_BIN)
    return CLI_BIN


def build_synth_lib() -> str:
    src = os.path.join(PKG_DIR, "synth", "rrg_synth.c")
    tmp = SYNTH_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", src, "-o", tmp, "-lz", "-lpthread", "-lm"])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


_gpu = None
_synth = None


def gpu_lib() -> ctypes.CDLL:
    """The decoder library (RRG_GPU_LIB overrides the in-tree path)."""
    global _gpu
    if _gpu is None:
        path = os.environ.get("RRG_GPU_LIB", GPU_LIB)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
        _gpu = ctypes.CDLL(path)
    return _gpu


def synth_lib() -> ctypes.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_LIB):
            raise RuntimeError(f"{SYNTH_LIB} is not built: run __graft_entry__.build()")
        _synth = ctypes.CDLL(SYNTH_LIB)
    return _synth
