"""The Fortran driver (mglc_amd/bin/mglc_cavity) end to end on the GPU: the
reference-style program reads a namelist, runs the loop of main.f90:85-103
through the C-ABI, and writes the reference's output files.  Its raw fields
must match the oracle within the FP64 parity bar."""
import os
import subprocess

import numpy as np
import pytest

from mglc_amd import build as B

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_fortran_driver_matches_oracle(tmp_path, oracle_lib):
    exe = os.path.join(B.BIN_DIR, "mglc_cavity")
    assert os.path.exists(exe), "Fortran driver not built"
    nx, ny, nz, itc_max, Re = 24, 20, 18, 599, 50.0
    nml = tmp_path / "run.nml"
    nml.write_text(f"&cavity\n total_nx={nx}, total_ny={ny}, total_nz={nz}, Reynolds={Re}, U0=0.1,\n"
                   f" itc_max={itc_max}, eps=0.0, check_every=200, collision_model='MRT',\n"
                   f" output_dir='{tmp_path}', write_raw=.true., write_binary=.true. /\n")
    print(f"[gpu] mglc_cavity {nx}x{ny}x{nz} itc_max={itc_max}", flush=True)
    r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    steps = itc_max + 1  # the reference loop runs itc = 1 .. itc_max + 1 (main.f90:85-87)
    assert "Successfully" in r.stdout
    raw = np.fromfile(tmp_path / f"cavity-raw-{steps:09d}.bin", dtype=np.float64)
    n = nx * ny * nz
    got = [raw[q * n:(q + 1) * n].reshape(nz, ny, nx) for q in range(4)]
    O = oracle_lib
    tau = 0.1 * nx / Re * 3.0 + 0.5
    rho, u, v, w, errs = O.cavity_run(O.cavity_params(nx, ny, nz, tau, 0.1), steps, check_every=200)
    for g, ref, name in zip(got, (rho, u, v, w), ("rho", "u", "v", "w")):
        assert relmax(g, ref) <= 1e-12, name
    # the residual lines itc errorU at 200, 400, 600 match the oracle's check()
    lines = [l.split() for l in r.stdout.splitlines() if len(l.split()) == 2 and l.split()[0].isdigit()]
    got_err = [float(e) for _, e in lines]
    np.testing.assert_allclose(got_err, errs, rtol=1e-9)
    # reference output files exist with the expected sizes
    plt = tmp_path / f"MRTcavity-{steps:09d}.plt"
    assert plt.exists() and plt.stat().st_size > 7 * 4 * n
    uz = np.loadtxt(tmp_path / f"u-z_{steps:09d}.dat")
    assert uz.shape == (nz, 2)
    np.testing.assert_allclose(uz[:, 0], u[:, (ny - 1) // 2, (nx - 1) // 2] / 0.1, rtol=1e-10, atol=1e-13)


def test_fortran_driver_restart_continues_bit_for_bit(tmp_path):
    """backup_file / restart_file (the reference's backupData / loadInitField):
    400 steps, save, restart to 800 equals one straight run of 800 steps."""
    exe = os.path.join(B.BIN_DIR, "mglc_cavity")
    nx, ny, nz = 20, 18, 16
    common = (f" total_nx={nx}, total_ny={ny}, total_nz={nz}, Reynolds=80.0, U0=0.1, eps=0.0,"
              f" check_every=200, storage='F32', write_tecplot=.false., write_centerlines=.false.,"
              f" write_raw=.true.,")
    runs = {
        "straight": f"itc_max=799, output_dir='{tmp_path}/straight'",
        "first": f"itc_max=399, output_dir='{tmp_path}/first', backup_file='{tmp_path}/state.bin'",
        "second": f"itc_max=799, output_dir='{tmp_path}/second', restart_file='{tmp_path}/state.bin'",
    }
    out = {}
    for name, extra in runs.items():
        os.makedirs(tmp_path / name, exist_ok=True)
        nml = tmp_path / f"{name}.nml"
        nml.write_text(f"&cavity\n{common}\n {extra} /\n")
        print(f"[gpu] mglc_cavity restart test: {name}", flush=True)
        r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr
        out[name] = r.stdout
    a = np.fromfile(tmp_path / "straight" / "cavity-raw-000000800.bin", dtype=np.float64)
    b = np.fromfile(tmp_path / "second" / "cavity-raw-000000800.bin", dtype=np.float64)
    assert a.size == 4 * nx * ny * nz and np.array_equal(a, b)
    assert "restarted from" in out["second"]

    def residuals(text):
        return {int(p[0]): p[1] for p in (l.split() for l in text.splitlines()) if len(p) == 2 and p[0].isdigit()}

    ra, rb = residuals(out["straight"]), residuals(out["second"])
    assert ra[600] == rb[600] and ra[800] == rb[800]
