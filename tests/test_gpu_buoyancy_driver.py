"""The buoyancy-cavity Fortran driver (mglc_amd/bin/mglc_buoyancy) end to end
on the GPU: it reads a &buoyancy namelist, runs the loop of the reference's
main.F90 (Buoyancy_driven_cavity/fortran/3d/mpi_blocked/main.F90:120-150)
through the C-ABI and writes the reference's output files.  Its raw fields
equal the Python path's bit for bit and the oracle's within the FP64 bar;
its calNuRe series and stationarity / convergence times equal those of
mglc_amd.statistics on the same values."""
import os
import subprocess

import numpy as np
import pytest

from mglc_amd import build as B
from mglc_amd.lbm import Cavity, ThermalConfig, buoyancy_initial_temperature
from mglc_amd.statistics import NuReStatistics

pytestmark = pytest.mark.gpu

EXE = os.path.join(B.BIN_DIR, "mglc_buoyancy")


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_driver(tmp_path, body, name="run"):
    nml = tmp_path / f"{name}.nml"
    nml.write_text(f"&buoyancy\n {body}, output_dir='{tmp_path}' /\n")
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "Successfully" in r.stdout
    return r.stdout


def test_buoyancy_driver_matches_python_path_and_oracle(tmp_path, oracle_lib):
    assert os.path.exists(EXE), "Fortran driver not built"
    nx, ny, nz, steps, every = 20, 18, 16, 600, 200
    Ra, Ek, seed, amp = 1e5, 0.01, 2024, 1e-4
    print(f"[gpu] mglc_buoyancy {nx}x{ny}x{nz} steps={steps} Ekman={Ek}", flush=True)
    out = run_driver(tmp_path, (
        f"total_nx={nx}, total_ny={ny}, total_nz={nz}, Rayleigh={Ra}, Prandtl=0.71, Mach=0.1,"
        f" Thot=0.5, Tcold=-0.5, Ekman={Ek}, epsU=0.0, epsT=0.0, check_every={every},"
        f" itc_limit={steps - 1}, noise_amplitude={amp}, noise_seed={seed},"
        f" write_raw=.true., write_binary=.true., write_tecplot=.true."))
    n = nx * ny * nz
    raw = np.fromfile(tmp_path / f"buoyancy-raw-{steps:09d}.bin", dtype=np.float64)
    assert raw.size == 5 * n
    got = [raw[q * n:(q + 1) * n].reshape(nz, ny, nx) for q in range(5)]

    # the same run through the Python binding: identical kernels and call split
    cfg = ThermalConfig(nx, ny, nz, Rayleigh=Ra, Prandtl=0.71, Mach=0.1, Thot=0.5, Tcold=-0.5, Ekman=Ek)
    T0 = buoyancy_initial_temperature(nx, ny, nz, t_hot=0.5, t_cold=-0.5, seed=seed, amp=amp)
    z = np.zeros_like(T0)
    fields = (np.ones_like(T0), z, z, z, T0)
    errs = []
    with Cavity(cfg) as c:
        c.load_fields(*fields)
        for _ in range(steps // every):
            c.step(every)
            errs.append(c.check_thermal())
        py = (*c.macro(), c.temperature())
    for g, p, name in zip(got, py, ("rho", "u", "v", "w", "T")):
        assert np.array_equal(g, p), name

    # against the oracle (FP64 parity bar)
    O = oracle_lib
    d = cfg.derived()
    p = O.thermal_params(nx, ny, nz, d["tau"], d["para_a"], d["g_beta"], d["omega_rot"], 0.5, -0.5, 0.0,
                         tconst_axis=1)
    ref = O.thermal_run(p, steps, fields=fields, check_every=every)
    U = float(np.sqrt(d["g_beta"] * nz))
    assert relmax(got[0], ref[0]) <= 1e-12
    assert relmax(got[4], ref[4]) <= 1e-12
    for g, r, name in zip(got[1:4], ref[1:4], "uvw"):
        assert np.max(np.abs(g - r)) <= 1e-12 * U, name
    # the check lines "itc errorU errorT" (check.F90:43-45)
    lines = [l.split() for l in out.splitlines() if len(l.split()) == 3 and l.split()[0].isdigit()]
    got_err = np.array([[float(a), float(b)] for _, a, b in lines])
    np.testing.assert_allclose(got_err, ref[5], rtol=1e-9)
    np.testing.assert_allclose(got_err, np.array(errs), rtol=0, atol=0)
    # output files of output.F90
    plt = tmp_path / f"buoyancyCavity-{steps}.plt"
    assert plt.exists() and plt.stat().st_size > 7 * 4 * n
    binf = tmp_path / f"buoyancyCavity-{steps}.bin"
    assert binf.stat().st_size == 4 * (8 * n + 8)   # four sequential records with their markers


def test_buoyancy_driver_statistics_and_stop(tmp_path):
    """statistics = .true.: calNuRe every int(timeUnit) steps; the run stops
    when statisticallyConvergeState is set (main.F90:120).  A weakly driven
    8^3 cavity settles within a few hundred time units; the driver's series,
    its stationary and converged times and its last step equal those of the
    Python statistics on the Python path's values."""
    nx = ny = nz = 8
    Ra, Ma = 2000.0, 0.3
    print(f"[gpu] mglc_buoyancy statistics {nx}^3 Ra={Ra} Ma={Ma}", flush=True)
    out = run_driver(tmp_path, (
        f"total_nx={nx}, total_ny={ny}, total_nz={nz}, Rayleigh={Ra}, Prandtl=0.71, Mach={Ma},"
        f" Thot=0.5, Tcold=-0.5, Ekman=0.0, epsU=0.0, epsT=0.0, dimensionlessTimeMax=2000,"
        f" statistics=.true., write_tecplot=.false."))
    cfg = ThermalConfig(nx, ny, nz, Rayleigh=Ra, Prandtl=0.71, Mach=Ma, Thot=0.5, Tcold=-0.5)
    d = cfg.derived()
    unit = int(np.sqrt(nz / d["g_beta"]))
    nu_f = np.loadtxt(tmp_path / "Nu_Bulk.dat")
    re_f = np.loadtxt(tmp_path / "Re_Bulk.dat")
    count = nu_f.shape[0]
    assert re_f.shape[0] == count

    # the same series from the Python binding, fed to the Python statistics
    stats = NuReStatistics()
    with Cavity(cfg) as c:
        c.initial()
        for t in range(count):
            c.step(unit)
            stats.add(c.nusselt()[0], c.reynolds())
            if stats.converged:
                break
    assert stats.time == count
    np.testing.assert_allclose(nu_f[:, 1], stats.nu, rtol=1e-12)
    np.testing.assert_allclose(re_f[:, 1], stats.re, rtol=1e-12)
    np.testing.assert_allclose(nu_f[:, 0], np.arange(1, count + 1) * unit / np.sqrt(nz / d["g_beta"]), rtol=1e-12)
    assert stats.stationary and stats.converged
    txt = (tmp_path / "statistically.txt").read_text()

    def time_after(marker, text):
        line = next(l for l in text.splitlines() if marker in l)
        return int(line.split("=")[-1].split()[0]) if "=" in line else int(line.split()[-1])

    assert time_after("Stationary Reached at Time", txt) == stats.stationary_time
    assert time_after("Converge Reached at Time", txt) == count
    assert time_after("statistically converged at time", out) == count
