"""Build the native library in-tree.

    python -m spherocyls_amd.build

produces ``spherocyls_amd/libspherocyls.so`` (HIP kernels for gfx950 + the C ABI
of include/spherocyls.h + the C++ SpherocylBox host class) and the command-line
drivers under ``spherocyls_amd/bin/``.  hipcc cross-compiles gfx950 code
without a GPU, so this runs on any machine with ROCm.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libspherocyls.so")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPHEROCYLS_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include")]

HIP_SOURCES = ["sc_device.hip", "sc_list.hip", "sc_api.hip", "sc_mpnn.hip", "sc_noalloc.hip"]
CXX_SOURCES = [os.path.join(CSRC, "sc_configgen.cpp")]
HOST_LIB_SOURCES = ["spherocyl_box.cpp", "dat_file.cpp"]
DRIVERS = {
    "spherocyl_strain": ("main_strain.cpp", []),
    "spherocyl_box": ("main_box.cpp", []),
    "spherocyl_pureshear": ("main_box.cpp", ["-DSPHEROCYL_PURE_SHEAR"]),
    "spherocyl_resize": ("main_resize.cpp", []),
    "spherocyl_minimize": ("main_relax.cpp", []),
    "spherocyl_qscompress": ("main_qs2p.cpp", []),
    "spherocyl_findjam": ("main_qs2p.cpp", ["-DSPHEROCYL_FIND_JAM"]),
    "spherocyl_ycompress": ("main_qs2p.cpp", ["-DSPHEROCYL_Y_COMPRESS"]),
    "spherocyl_findjam_1p": ("main_qs2p.cpp", ["-DSPHEROCYL_FIND_JAM_1P"]),
    "spherocyl_qscycle": ("main_qs2p.cpp", ["-DSPHEROCYL_QS_CYCLE"]),
}


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    host_srcs = list(CXX_SOURCES)
    host_srcs += [os.path.join(HOST, f) for f in HOST_LIB_SOURCES
                  if os.path.exists(os.path.join(HOST, f))]
    for src in host_srcs:
        obj = os.path.join(BUILD, os.path.basename(src).replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", src, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for exe, (src, defs) in DRIVERS.items():
        path = os.path.join(HOST, src)
        if not os.path.exists(path):
            continue
        _run(["g++", *CXX_FLAGS, *defs, path, "-o", os.path.join(BIN, exe), f"-L{PKG}", "-lspherocyls",
              f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
