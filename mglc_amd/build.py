"""Build steps for the native parts: the HIP library (hipcc, gfx950) and the
Fortran driver (amdflang).  Outputs stay in the source tree so they travel
with it."""
from __future__ import annotations

import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmglc_lbm.so")
FORTRAN_DIR = os.path.join(HERE, "fortran")
BIN_DIR = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("MGLC_FC", "/opt/rocm/bin/amdflang")

# -packed-fp32-ops: no v_pk_*_f32 in generated code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass reports the device-only feature flag as unknown; drop that line
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({' '.join(cmd)}):\n{out}")
    return out


def build_library() -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran_driver() -> list[str]:
    """The Fortran programs that keep the reference's parameter modules and
    loops and call the library through iso_c_binding: mglc_cavity (the
    lid-driven cavity) and mglc_buoyancy (the buoyancy-driven cavity)."""
    os.makedirs(BIN_DIR, exist_ok=True)
    mod_dir = os.path.join(BIN_DIR, "mod")
    os.makedirs(mod_dir, exist_ok=True)
    programs = {
        "mglc_cavity": ["mglc_lbm_capi.f90", "commondata.f90", "output.f90", "main.f90"],
        "mglc_buoyancy": ["mglc_lbm_capi.f90", "buoyancy_data.f90", "output.f90", "buoyancy_main.f90"],
    }
    exes = []
    for name, order in programs.items():
        srcs = [os.path.join(FORTRAN_DIR, f) for f in order]
        exe = os.path.join(BIN_DIR, name)
        _run([FC, "-O2", "-J", mod_dir, *srcs, "-o", exe, f"-L{HERE}", "-lmglc_lbm",
              f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."])
        exes.append(exe)
    return exes
