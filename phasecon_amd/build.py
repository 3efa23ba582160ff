"""Builds the in-tree native libraries for gfx950.

  phasecon_amd/libmbir4d.so  -- HIP kernels (csrc/*.hip, hipcc) + the C host
                                engine (csrc/*.c, gcc), exporting reconstruct()
                                and the C-ABI of include/mbir_hip.h
  phasecon_amd/XT_Main       -- the command-line driver with the reference's
                                flags (reconstruct/basic/XT_Main.c)

Usage: python -m phasecon_amd.build   (always recompiles)
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INC,
    # packed FP32 instructions stay off (they are not trusted on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-Wall", "-I", INC,
           "-I", os.path.join(ROCM, "include")]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hip_sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def c_sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".c") and f != "xt_main.c")


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in hip_sources():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in c_sources():
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(PKG, "libmbir4d.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lm"])
    os.replace(tmp, lib)
    main_src = os.path.join(CSRC, "xt_main.c")
    if os.path.exists(main_src):
        exe = os.path.join(PKG, "XT_Main")
        _run(["gcc", *C_FLAGS, "-o", exe, main_src, f"-L{PKG}", "-lmbir4d",
              f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + PKG, "-lm"])
    return lib


if __name__ == "__main__":
    build()
    sys.exit(0)
