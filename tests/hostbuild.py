"""Builds the host programs' CPU test variants into tests/emu/_build (test infrastructure).

The slave and the standalone aligner are linked once against libsxc_hip.so (the product,
built by satsuma2_amd/build.py) and once against the workgroup emulator's copy of the same
C ABI (tests/emu/sxc_emu.cpp, SXC_EMU_ABI), so their flag handling, file handling and master
exchange run on a machine without a GPU while computing with the kernels' own code.
"""
from __future__ import annotations

import os
import subprocess

from conftest import ROOT

CSRC = os.path.join(ROOT, "satsuma2_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "emu")
BUILD = os.path.join(EMU, "_build")
LIBDIR = os.path.join(ROOT, "satsuma2_amd", "lib")

KERNEL_DEPS = [os.path.join(CSRC, f) for f in ("sxc_fft.h", "sxc_pipeline.h", "sxc_scan.h", "sxc_types.h",
                                               "sxc_batch_host.h", "sxc_tables.h", "sxc_tables.cpp")]
HOST_DEPS = [os.path.join(CSRC, "host", f) for f in sorted(os.listdir(os.path.join(CSRC, "host")))]
PROGRAM_SOURCES = {
    "HomologyByXCorrSlave": ["host/slave_main.cpp", "host/sxc_seqio.cpp"],
    "HomologyByXCorr": ["host/xcorr_main.cpp", "host/sxc_seqio.cpp"],
}


def _stale(target, deps):
    return not os.path.exists(target) or os.path.getmtime(target) < max(os.path.getmtime(d) for d in deps)


def emu_program(name: str, sanitize: bool = False) -> str:
    """The program `name` linked against the emulator (optionally with AddressSanitizer)."""
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, name + ("_asan" if sanitize else "") + "_emu")
    srcs = [os.path.join(CSRC, s) for s in PROGRAM_SOURCES[name]] + [
        os.path.join(EMU, "sxc_emu.cpp"), os.path.join(CSRC, "sxc_tables.cpp")]
    if _stale(exe, srcs + KERNEL_DEPS + HOST_DEPS):
        opt = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"] if sanitize else ["-O2"]
        subprocess.run(["g++", *opt, "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-DSXC_EMU_ABI",
                        "-pthread", "-o", exe + ".tmp", *srcs], check=True)
        os.replace(exe + ".tmp", exe)
    return exe


def seqio_check() -> str:
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "seqio_check")
    srcs = [os.path.join(EMU, "seqio_check.cpp"), os.path.join(CSRC, "host", "sxc_seqio.cpp")]
    if _stale(exe, srcs + HOST_DEPS):
        subprocess.run(["g++", "-O1", "-g", "-fsanitize=address,undefined", "-std=c++17", "-o", exe + ".tmp", *srcs],
                       check=True)
        os.replace(exe + ".tmp", exe)
    return exe


def product_program(name: str) -> str:
    """The program as built by satsuma2_amd/build.py (linked against libsxc_hip.so)."""
    exe = os.path.join(LIBDIR, name)
    if not os.path.exists(exe):
        raise FileNotFoundError(f"{exe} is not built: run python -m satsuma2_amd.build")
    return exe
