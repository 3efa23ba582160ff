"""Builds the in-tree native libraries for gfx950.

  phasecon_amd/libmbir4d.so    -- MBIR: HIP kernels (csrc/mbir_hip.hip) + the C
                                  host engine (csrc/mbir_engine.c), exporting
                                  reconstruct() and the C-ABI of
                                  include/mbir_hip.h; the learned warm start
                                  (csrc/unet_hip.hip, include/unet_hip.h)
  phasecon_amd/libpcmbir4d.so  -- PCMBIR phase retrieval: csrc/pcmbir_*.hip /
                                  pcmbir_*.c, the C-ABI of include/pcmbir_hip.h

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


# library -> source files in csrc/ (prefix match)
LIBS = {
    "libmbir4d.so": ("mbir_", "unet_"),
    "libpcmbir4d.so": ("pcmbir_",),
}


# extra link arguments: the phase-contrast simulator projects its phantoms
# with the MBIR library's A rule (built first, found next to this one)
LINK = {
    "libpcmbir4d.so": ["-L", PKG, "-lmbir4d", "-Wl,-rpath,$ORIGIN"],
}


def sources(prefixes):
    return sorted(f for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".c")) and f.startswith(prefixes))


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    out = []
    for lib_name, prefixes in LIBS.items():
        objs = []
        for src in sources(prefixes):
            obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
            if src.endswith(".hip"):
                _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            else:
                _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        lib = os.path.join(PKG, lib_name)
        tmp = lib + ".tmp"
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
              *LINK.get(lib_name, []), "-lm"])
        os.replace(tmp, lib)
        out.append(lib)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
