"""Builds the in-tree native libraries for gfx950.

  phasecon_amd/libmbir4d.so    -- MBIR: HIP kernels (csrc/mbir_hip.hip) + the C
                                  host engine (csrc/mbir_engine.c), exporting
                                  reconstruct() and the C-ABI of
                                  include/mbir_hip.h; the learned warm start
                                  (csrc/unet_hip.hip, include/unet_hip.h)
  phasecon_amd/libpcmbir4d.so  -- PCMBIR phase retrieval: csrc/pcmbir_*.hip /
                                  pcmbir_*.c, the C-ABI of include/pcmbir_hip.h
  phasecon_amd/libmbir4d_hdf5.so -- the real-data drivers' HDF5 input stages
                                  (csrc/mbir_hdf5.c, include/mbir_hdf5.h) over
                                  libmbir4d; built where libhdf5 is found

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


# library -> source files in csrc/ (prefix match, less the files of later
# libraries listed in OWN)
LIBS = {
    "libmbir4d.so": ("mbir_", "unet_"),
    "libpcmbir4d.so": ("pcmbir_",),
    "libmbir4d_hdf5.so": ("mbir_hdf5",),
}
OWN = {"mbir_hdf5.c": "libmbir4d_hdf5.so"}
HDF5 = os.environ.get("HDF5_ROOT", "/opt/conda")


# extra link arguments: the phase-contrast simulator projects its phantoms
# with the MBIR library's A rule (built first, found next to this one)
LINK = {
    "libpcmbir4d.so": ["-L", PKG, "-lmbir4d", "-Wl,-rpath,$ORIGIN"],
    "libmbir4d_hdf5.so": ["-L", PKG, "-lmbir4d", "-Wl,-rpath,$ORIGIN",
                          "-L", os.path.join(HDF5, "lib"), "-lhdf5",
                          "-Wl,--enable-new-dtags,-rpath," + os.path.join(HDF5, "lib")],
}
EXTRA_INC = {"libmbir4d_hdf5.so": ["-I", os.path.join(HDF5, "include")]}


def sources(lib_name, prefixes):
    return sorted(f for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".c")) and f.startswith(prefixes)
                  and OWN.get(f, lib_name) == lib_name)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    out = []
    for lib_name, prefixes in LIBS.items():
        if lib_name == "libmbir4d_hdf5.so" and not os.path.exists(
                os.path.join(HDF5, "include", "hdf5.h")):
            print(f"skipping {lib_name}: no hdf5.h under {HDF5}", flush=True)
            continue
        objs = []
        for src in sources(lib_name, prefixes):
            obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
            if src.endswith(".hip"):
                _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            else:
                _run(["gcc", *C_FLAGS, *EXTRA_INC.get(lib_name, []), "-c",
                      os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        lib = os.path.join(PKG, lib_name)
        tmp = lib + ".tmp"
        if any(o.endswith(".o") and os.path.basename(o).startswith("mbir_hdf5") for o in objs):
            # plain C over libmbir4d: linked by gcc, so that libhdf5's
            # directory serves only libhdf5 (not the HIP runtime's libstdc++)
            _run(["gcc", "-shared", "-fPIC", "-o", tmp, *objs, *LINK.get(lib_name, []), "-lm"])
        else:
            _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
                  *LINK.get(lib_name, []), "-lm"])
        os.replace(tmp, lib)
        out.append(lib)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
