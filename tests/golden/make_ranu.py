"""Writes tests/golden/ranu_series.npz: the reference's own Nusselt /
Reynolds statistics (calNuRe, Buoyancy_driven_cavity/fortran/3d/mpi_blocked/
RaNu.F90, built unmodified with its module.F90 into oracle/_ref/ref_ranu by
oracle/ref/Makefile) run over 960 time units of given fields, and what it
computed and wrote.

Fields on a 4 x 3 x 2 grid.  The first three time units hold seeded random
fields (they pin the Nu and Re formulas); the rest hold uniform fields
(w = a, T = 1, u = b, v = 0) chosen so that Nu and Re follow a decaying
oscillation with seeded noise, a late bump in both, and so reach
stationarity at time 500, fail the first convergence test (700) and pass
the second (900): every branch of calNuRe without flow reversal (the
reference's compile-time flowReversal = 0).

Run in this container, where the reference sources are: python tests/golden/make_ranu.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_ranu")
SHAPE = (2, 3, 4)  # (nz, ny, nx)
NTIMES, ITC_STEP, SEED = 960, 1000, 11


def run(fields, workdir):
    """fields: (ntimes, 4, nz, ny, nx) of u, v, w, T; returns (diffusivity,
    viscosity, timeUnit), the (ntimes, 7) records and the three files."""
    nz, ny, nx = SHAPE
    fin, fout = os.path.join(workdir, "fields.bin"), os.path.join(workdir, "out.bin")
    np.ascontiguousarray(fields, dtype=np.float64).tofile(fin)
    subprocess.run([EXE, str(nx), str(ny), str(nz), str(len(fields)), str(ITC_STEP), fin, fout], cwd=workdir,
                   check=True, stdout=subprocess.DEVNULL)
    raw = np.fromfile(fout, dtype=np.float64)
    files = {}
    for name in ("Nu_Bulk.dat", "Re_Bulk.dat", "statistically.txt"):
        path = os.path.join(workdir, name)
        files[name] = open(path, "rb").read() if os.path.exists(path) else b""
    return raw[:3], raw[3:].reshape(len(fields), 7), files


def targets(n):
    rng = np.random.default_rng(SEED)
    t = np.arange(1, n + 1)
    nu = (4.6 + 1.5 * np.exp(-t / 70.0) * np.cos(t / 9.0) + 0.05 * np.exp(-t / 300.0)
          + 0.004 * rng.standard_normal(n) + 0.02 * np.exp(-((t - 430) / 40.0) ** 2)
          + 0.12 * np.exp(-((t - 660) / 30.0) ** 2))
    re = 31.0 + 9.0 * np.exp(-t / 90.0) + 0.05 * rng.standard_normal(n) + 0.6 * np.exp(-((t - 430) / 40.0) ** 2)
    return nu, re


def make_fields(kappa, nu_visc):
    nz = SHAPE[0]
    nu, re = targets(NTIMES)
    fields = np.zeros((NTIMES, 4) + SHAPE)
    rng = np.random.default_rng(SEED + 1)
    for q in range(NTIMES):
        if q < 3:
            fields[q, :3] = rng.uniform(-0.05, 0.05, (3,) + SHAPE)
            fields[q, 3] = rng.uniform(-0.5, 0.5, SHAPE)
            continue
        a = (nu[q] - 1.0) * kappa / nz
        b = np.sqrt((re[q] * nu_visc / nz) ** 2 - a * a)
        fields[q, 0], fields[q, 2], fields[q, 3] = b, a, 1.0
    return fields


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} is not built (make -f oracle/ref/Makefile)")
    with tempfile.TemporaryDirectory() as d:
        consts, _, _ = run(np.zeros((1, 4) + SHAPE), d)
    kappa, nu_visc, time_unit = consts
    fields = make_fields(kappa, nu_visc)
    with tempfile.TemporaryDirectory() as d:
        consts, rec, files = run(fields, d)
    out = os.path.join(HERE, "ranu_series.npz")
    np.savez_compressed(out, fields=fields, consts=consts, rec=rec, itc_step=ITC_STEP,
                        nu_bulk=np.frombuffer(files["Nu_Bulk.dat"], dtype=np.uint8),
                        re_bulk=np.frombuffer(files["Re_Bulk.dat"], dtype=np.uint8),
                        statistically=np.frombuffer(files["statistically.txt"], dtype=np.uint8))
    st = rec[:, 4:7]
    print(f"wrote {out}: kappa {kappa}, nu {nu_visc}, timeUnit {time_unit}; stationary from time "
          f"{int(st[-1, 1])}, converged: {int(st[-1, 2])} (first at time {int(np.argmax(st[:, 2] > 0)) + 1})")


if __name__ == "__main__":
    main()
