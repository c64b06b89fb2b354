"""Generate the golden fixtures of the D3Q19 cavity and the buoyancy cavity
from the reference itself.

Runs oracle/_ref/ref_cavity -- the reference's own Fortran subroutines
(MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/{commondata,initial,
collision,streaming,bounce_back,macro}.f90), compiled unmodified by
oracle/ref/Makefile -- and stores rho, u, v, w after n steps.  Runs only in a
container that has /root/reference; the fixtures are committed.

The reference's tauf, U0 and Reynolds are compile-time parameters
(commondata.f90:4-9: Re = 1000, U0 = 0.1, N = 65, tauf = 0.5195), so the
fixtures are MRT at that tau on small grids of varying shape.

Buoyancy fixtures run oracle/_ref/ref_thermal, the reference's own
module/initial/evolution_f/evolution_g.F90 built with its Makefile's macros
(no-slip walls, y walls held at Thot = 1 / Tcold = 0, other walls adiabatic),
at its compile-time parameters (module.F90: Ra = 1e6, Pr = 0.71, Ma = 0.1,
N = 51, Ekman = 0.001 so the Coriolis force is on).

Moving-disc fixtures run oracle/_ref/ref_dkt, the reference's sequential
two-disc program MPI/Micro_particles/fortran/case4/seq/CaseIV_DKT.F90 compiled
unmodified (201 x 801 channel, two discs of radius 10 falling in a moving
window), and keep after n steps the disc state (all 19 scalars), the sha256
of each whole field (rho, u, v, obst as (801, 201); f as (801, 201, 9),
float64 little-endian), and rho, u, v in a 24 x 24 window at each disc.  At
2400 steps the discs are 21.9 apart, inside the 24 at which the
particle-particle spring acts (from step 2245 on).

usage: python tests/golden/make_golden.py [cavity] [thermal] [dkt]   (default: all)
"""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

# (nx, ny, nz, steps): shapes are non-cubic so that a transposed axis fails.
CASES = [
    (8, 9, 10, 1),
    (8, 9, 10, 2),
    (8, 9, 10, 10),
    (8, 9, 10, 100),
    (12, 10, 9, 300),
    (16, 16, 16, 500),
    (20, 17, 13, 1000),
    # edge shapes: the smallest grid (every node touches a wall and the lid),
    # two-node-thin slabs along each axis, and x extents around one and two
    # wavefronts (65: a one-lane tail; 33: half a wave)
    (2, 2, 2, 50),
    (2, 7, 3, 40),
    (3, 2, 5, 60),
    (4, 5, 2, 60),
    (33, 17, 5, 200),
    (65, 3, 4, 100),
]

THERMAL_CASES = [
    (8, 9, 10, 1),
    (8, 9, 10, 10),
    (12, 10, 9, 300),
    (16, 16, 16, 500),
    (20, 17, 13, 1000),
    # edge shapes: three nodes between the held walls, two-node-thin slabs,
    # an x extent past one wavefront
    (3, 3, 3, 60),
    (2, 5, 4, 80),
    (5, 4, 2, 80),
    (65, 6, 5, 100),
]

DKT_STEPS = [1, 300, 2400]
DKT_WINDOW = 24


def dkt_window(state, c):
    """Rows and columns of the DKT_WINDOW-wide window at disc c, in the
    (801, 201) field layout (1-based node i, j at [j - 1, i - 1])."""
    ny, nx = state["rho"].shape
    h = DKT_WINDOW // 2
    i0 = min(max(int(round(state["xc"][c])) - h, 0), nx - DKT_WINDOW)
    j0 = min(max(int(round(state["yc"][c])) - h, 0), ny - DKT_WINDOW)
    return slice(j0, j0 + DKT_WINDOW), slice(i0, i0 + DKT_WINDOW)


def dkt_digest(state):
    """sha256 of each whole field of a DKT state."""
    return {k: hashlib.sha256(np.ascontiguousarray(state[k], dtype="<f8").tobytes()).hexdigest()
            for k in ("rho", "u", "v", "obst", "f")}


def write_dkt():
    for ns in DKT_STEPS:
        st = O.run_reference_dkt(ns)
        scal = np.array([st["itc"], st["shift"], st["rho_avg"]] +
                        [x for name in O.DKT_SCALARS[3:] for x in st[name]], dtype=np.float64)
        win = {}
        for c in range(2):
            js, is_ = dkt_window(st, c)
            for k in ("rho", "u", "v"):
                win[f"{k}{c}"] = st[k][js, is_]
            win[f"origin{c}"] = np.array([js.start, is_.start])
        dig = dkt_digest(st)
        name = os.path.join(HERE, f"dkt_s{ns}.npz")
        np.savez_compressed(name, steps=ns, scal=scal, **{f"sha_{k}": np.array(v) for k, v in dig.items()}, **win)
        print("wrote", os.path.relpath(name, ROOT))


def main():
    which = set(sys.argv[1:]) or {"cavity", "thermal", "dkt"}
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref", "Makefile")], check=True)
    if "dkt" in which:
        write_dkt()
    with tempfile.TemporaryDirectory() as td:
        for nx, ny, nz, ns in CASES if "cavity" in which else []:
            hdr, rho, u, v, w = O.run_reference_cavity(nx, ny, nz, ns, os.path.join(td, "out.bin"))
            name = os.path.join(HERE, f"cavity_mrt_{nx}x{ny}x{nz}_s{ns}.npz")
            np.savez_compressed(name, nx=nx, ny=ny, nz=nz, steps=ns, tauf=hdr["tauf"], U0=hdr["U0"],
                                rho0=hdr["rho0"], rho=rho, u=u, v=v, w=w)
            print("wrote", os.path.relpath(name, ROOT))
        for nx, ny, nz, ns in THERMAL_CASES if "thermal" in which else []:
            hdr, (rho, u, v, w, T) = O.run_reference_thermal(nx, ny, nz, ns, os.path.join(td, "out.bin"))
            name = os.path.join(HERE, f"thermal_{nx}x{ny}x{nz}_s{ns}.npz")
            np.savez_compressed(name, nx=nx, ny=ny, nz=nz, steps=ns, rho=rho, u=u, v=v, w=w, T=T,
                                **{k: float(val) for k, val in hdr.items()})
            print("wrote", os.path.relpath(name, ROOT))


if __name__ == "__main__":
    main()
