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


FLAGS = ["-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas"]
OPT = {False: ["-O2", "-fPIC"], True: ["-O1", "-fsanitize=address", "-fno-omit-frame-pointer"]}


def emu_objects(sanitize: bool = False) -> list:
    """The emulator (kernel bodies + its C ABI) and the host tables, compiled once per flavour."""
    os.makedirs(BUILD, exist_ok=True)
    tag = "asan" if sanitize else "o2"
    objs = []
    for src in (os.path.join(EMU, "sxc_emu.cpp"), os.path.join(CSRC, "sxc_tables.cpp")):
        obj = os.path.join(BUILD, f"{os.path.basename(src)}.{tag}.o")
        if _stale(obj, [src] + KERNEL_DEPS + [os.path.join(ROOT, "include", "sxc.h")]):
            subprocess.run(["g++", *OPT[sanitize], *FLAGS, "-c", "-o", obj + ".tmp", src], check=True)
            os.replace(obj + ".tmp", obj)
        objs.append(obj)
    return objs


def _link(exe, srcs, objs, sanitize):
    if _stale(exe, srcs + objs + HOST_DEPS):
        subprocess.run(["g++", *OPT[sanitize], *FLAGS, "-pthread", "-o", exe + ".tmp", *srcs, *objs], check=True)
        os.replace(exe + ".tmp", exe)
    return exe


def emu_program(name: str, sanitize: bool = False) -> str:
    """The program `name` linked against the emulator (optionally with AddressSanitizer)."""
    exe = os.path.join(BUILD, name + ("_asan" if sanitize else "") + "_emu")
    srcs = [os.path.join(CSRC, s) for s in PROGRAM_SOURCES[name]]
    return _link(exe, srcs, emu_objects(sanitize), sanitize)


def emu_library() -> str:
    """libsxc_emu.so: the emulator's entry points for ctypes."""
    so = os.path.join(BUILD, "libsxc_emu.so")
    objs = emu_objects(False)
    if _stale(so, objs):
        subprocess.run(["g++", "-shared", "-o", so + ".tmp", *objs], check=True)
        os.replace(so + ".tmp", so)
    return so


def emu_checker() -> str:
    """tests/emu/emu_check.cpp against the AddressSanitizer emulator."""
    exe = os.path.join(BUILD, "emu_check")
    return _link(exe, [os.path.join(EMU, "emu_check.cpp")], emu_objects(True), True)


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
