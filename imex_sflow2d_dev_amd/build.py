"""Builds the HIP library libimexhip.so for gfx950 in-tree, and the Fortran
side of the boundary: the ISO_C_BINDING module imex_hip_binding
(fortran/imex_hip_binding.f90 -> _lib/imex_hip_binding.o, .mod) and the
Fortran host program imex_fdriver linked against the library.

hipcc flags: -O3, -ffp-contract=off (no fused multiply-add: the FP64 path
keeps the reference's rounding), packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libimexhip.so")
SRCS = [os.path.join(HERE, "csrc", f)
        for f in ("imex_hip.hip", "imex_output.hip", "imex_pore.hip", "imex_writers.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("IMEX_FC", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FDRIVER = os.path.join(LIBDIR, "imex_fdriver")
ARCH = os.environ.get("IMEX_OFFLOAD_ARCH", "gfx950")


# the phase-clock build: the same library with per-cell shader-clock readings
# of the update kernel (imex_physics.h, PhaseSlot), read by phase_clocks.py
LIB_PHASES = os.path.join(LIBDIR, "libimexhip_phases.so")


def compile_cmd(out: str = LIB, defines=()):
    return [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")] + \
        [f"-D{d}" for d in defines] + ["-o", out] + SRCS


def build_phase_clocks() -> str:
    """_lib/libimexhip_phases.so, the library with the phase clocks."""
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB_PHASES + ".tmp"
    r = subprocess.run(compile_cmd(tmp, ("IMEX_PHASE_CLOCKS",)), capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libimexhip_phases.so")
    os.replace(tmp, LIB_PHASES)
    return LIB_PHASES


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = compile_cmd(tmp)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libimexhip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


def build_fortran() -> str:
    """imex_hip_binding.o/.mod and imex_fdriver in _lib (rpath $ORIGIN)."""
    obj = os.path.join(LIBDIR, "imex_hip_binding.o")
    steps = [
        [FC, "-O2", "-fPIC", "-c", os.path.join(FDIR, "imex_hip_binding.f90"), "-J", LIBDIR,
         "-o", obj],
        [FC, "-O2", os.path.join(FDIR, "imex_fdriver.f90"), obj, "-I", LIBDIR, "-L", LIBDIR,
         "-limexhip", "-Wl,-rpath,$ORIGIN", "-o", FDRIVER + ".tmp"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=LIBDIR)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("amdflang failed: " + " ".join(cmd))
    os.replace(FDRIVER + ".tmp", FDRIVER)
    return FDRIVER


if __name__ == "__main__":
    if "--phase-clocks" in sys.argv[1:]:
        print(build_phase_clocks())
    else:
        print(build(verbose=True))
