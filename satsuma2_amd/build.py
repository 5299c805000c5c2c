"""Builds libsxc_hip.so in-tree for gfx950 (hipcc, no JIT cache).

    python -m satsuma2_amd.build

Flags: -ffp-contract=off keeps every float/double operation separately rounded, as the
reference (plain -O3, no -march) computes them; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIBDIR, "libsxc_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SXC_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unknown-pragmas"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["sxc_kernels.hip", "sxc_kernels_r2.hip", "sxc_kernels_s32.hip", "sxc_kernels_occ.hip", "sxc_kernels_lean.hip",
           "sxc_api.hip",
           "sxc_server.hip", "sxc_static.hip", "sxc_tables.cpp", "sxc_kmatch.hip",
           # the master's side of the search (host C++): match sets, chaining, grid search
           "sxc_master_api.cpp", "host/sxc_master.cpp", "host/sxc_seqio.cpp"]
# Host programs linked against the library (plain C++, no device code).
PROGRAMS = {
    "HomologyByXCorrSlave": ["host/slave_main.cpp", "host/sxc_seqio.cpp"],
    "HomologyByXCorr": ["host/xcorr_main.cpp", "host/sxc_seqio.cpp"],
    # a caller of the reference-compatible C++ classes (include/satsuma_compat/)
    "compat_align": ["host/compat_align.cpp"],
    # kmatch/KMatch.cc's program with its k-mer tables on the GPU
    "KMatch": ["host/kmatch_main.cpp"],
}
CXX = os.environ.get("CXX", "g++")


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, src.replace("/", "_") + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if src.endswith(".hip") else []) + [
        "-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(verbose: bool = True) -> str:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    for name, srcs in PROGRAMS.items():
        exe = os.path.join(LIBDIR, name)
        cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-pthread", f"-I{os.path.join(os.path.dirname(PKG), 'include')}", "-o", exe + ".tmp",
               *[os.path.join(CSRC, s) for s in srcs], f"-L{LIBDIR}", "-lsxc_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        if verbose:
            print(f"built {exe}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
