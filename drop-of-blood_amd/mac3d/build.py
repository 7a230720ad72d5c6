"""Build recipe of the MAC step library (hipcc for gfx950, in-tree).

Outputs (git-ignored, shipped with the tree):
  drop-of-blood_amd/lib/libmac3d.so   HIP kernels + C ABI (include/mac3d.h)
  drop-of-blood_amd/lib/mac3d.node    Node N-API addon over the C ABI
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libmac3d.so")
NODE_ADDON = os.path.join(LIBDIR, "mac3d.node")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
NODE_INCLUDE = "/usr/include/node"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # contract products into FMAs as hipcc does by default, except in blocks
    # marked "#pragma clang fp contract(off)" (the hash ranks and R3 sequence,
    # which must round every product as the reference's f32 does)
    "-ffp-contract=fast-honor-pragmas",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "mac3d.hip"), "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_node_addon() -> str | None:
    """The N-API shim: plain C++ against node_api.h, linked to libmac3d.so."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        print("node_api.h not found; skipping the Node addon", flush=True)
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = NODE_ADDON + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          "-I", NODE_INCLUDE, "-I", os.path.join(ROOT, "include"),
          os.path.join(PKG, "napi", "mac3d_napi.cpp"), "-o", tmp,
          "-L", LIBDIR, "-lmac3d", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, NODE_ADDON)
    return NODE_ADDON


def build(node: bool = True) -> None:
    build_lib()
    if node and os.path.exists(os.path.join(PKG, "napi", "mac3d_napi.cpp")):
        build_node_addon()


if __name__ == "__main__":
    build(node="--no-node" not in sys.argv)
