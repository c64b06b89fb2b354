"""Writes tests/golden/output_files.json: the size and SHA-256 of each file
the reference's own output writers write for given fields:
  cavity    output_binary, output_Tecplot, getVelocity of
            MPI/Lid_driven_cavity/fortran/3d/seq/lid_driven_cavity_3d.f90
            (oracle/_ref/ref_output), fields() on its compile-time 65^3 grid;
  buoyancy  output_binary, output_Tecplot of
            MPI/Buoyancy_driven_cavity/fortran/3d/seq/bouyancy3d.F90
            (oracle/_ref/ref_output_buoyancy), fields_buoyancy() on its 51^3
            grid;
both built unmodified by oracle/ref/Makefile, both at itc = ITC.

The fields are integer hashes of the cell index scaled to flow-like
magnitudes, so any platform computes them bit for bit.

Run in this container, where the reference sources are: python tests/golden/make_output.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_output")
EXE_B = os.path.join(ROOT, "oracle", "_ref", "ref_output_buoyancy")
N, N_B, ITC, U0 = 65, 51, 123456, 0.1


def fields():
    """rho, u, v, w on N^3 (index x fastest)."""
    c = np.arange(N ** 3, dtype=np.int64)
    rho = 1.0 + ((c * 2654435761 + 7) % 20011 - 10005).astype(np.float64) * 1e-7
    u = ((c * 7919 + 13) % 4001 - 2000).astype(np.float64) * 5e-5
    v = ((c * 104729 + 5) % 3001 - 1500).astype(np.float64) * 3e-5
    w = ((c * 1299709 + 3) % 5003 - 2501).astype(np.float64) * 2e-5
    return rho, u, v, w


def fields_buoyancy():
    """u, v, w, T on N_B^3 (index x fastest)."""
    c = np.arange(N_B ** 3, dtype=np.int64)
    u = ((c * 7919 + 11) % 4001 - 2000).astype(np.float64) * 4e-6
    v = ((c * 104729 + 2) % 3001 - 1500).astype(np.float64) * 3e-6
    w = ((c * 1299709 + 9) % 5003 - 2501).astype(np.float64) * 5e-6
    T = ((c * 2654435761 + 1) % 20011 - 10005).astype(np.float64) * 5e-5
    return u, v, w, T


def run(exe, arrays):
    with tempfile.TemporaryDirectory() as d:
        np.concatenate(arrays).tofile(os.path.join(d, "fields.bin"))
        subprocess.run([exe, str(ITC)], cwd=d, check=True, stdout=subprocess.DEVNULL)
        return file_table(d)


def file_table(directory):
    out = {}
    for name in sorted(os.listdir(directory)):
        if name == "fields.bin":
            continue
        data = open(os.path.join(directory, name), "rb").read()
        out[name] = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
    return out


def main():
    for exe in (EXE, EXE_B):
        if not os.path.exists(exe):
            sys.exit(f"{exe} is not built (make -f oracle/ref/Makefile)")
    table = run(EXE, fields())
    table_b = run(EXE_B, fields_buoyancy())
    out = os.path.join(HERE, "output_files.json")
    with open(out, "w") as fh:
        json.dump({"n": N, "n_buoyancy": N_B, "itc": ITC, "U0": U0, "files": table, "files_buoyancy": table_b},
                  fh, indent=1)
        fh.write("\n")
    print(f"wrote {out}: {sorted(table)} {sorted(table_b)}")


if __name__ == "__main__":
    main()
