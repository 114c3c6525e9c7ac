"""Build the gfx950 shared library ``libbubbles_amd.so`` in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instructions disabled and FP contraction left to per-function pragmas (the
FP64 hash and bucket rule turn it off locally).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libbubbles_amd.so")

SOURCES = ["bbx_abi.hip", "bbx_grid.hip", "bbx_sph.hip", "bbx_mlp.hip", "bbx_stream.hip", "bbx_io.hip",
           "bbx_shapes.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
    "-I" + CSRC,
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


FACADE_DIR = os.path.join(ROOT, "tests", "cpp")
FACADE_BIN = os.path.join(FACADE_DIR, "facade_dam_break")
EMIT_BIN = os.path.join(FACADE_DIR, "facade_emit")
OBSTACLES_BIN = os.path.join(FACADE_DIR, "facade_obstacles")
LOADOBJ_BIN = os.path.join(FACADE_DIR, "facade_loadobj")


def build_facade_example() -> str:
    """Compile the C++ facade scenes (host code only) against the library."""
    cxx = shutil.which("g++") or "g++"
    for exe in (FACADE_BIN, EMIT_BIN, OBSTACLES_BIN, LOADOBJ_BIN):
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I" + INCLUDE, exe + ".cpp", "-o", exe,
               "-L" + PKG, "-lbubbles_amd", "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../bubbles_amd"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"{exe}.cpp failed to build:\n{res.stdout}\n{res.stderr}")
    return FACADE_BIN


if __name__ == "__main__":
    build(verbose=True)
    build_facade_example()
    sys.exit(0)
