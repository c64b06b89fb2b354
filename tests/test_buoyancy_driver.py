"""CPU tests of the buoyancy-cavity Fortran driver (mglc_amd/bin/mglc_buoyancy):
its parameter module derives the reference's constants
(Buoyancy_driven_cavity/fortran/3d/mpi_blocked/module.F90:19-33) as the Python
configuration does, and its SplitMix64 noise equals the Python generator bit
for bit.  `--initial-only` stops before any device call."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mglc_amd import build as B
from mglc_amd import lbm


def _driver():
    exe = os.path.join(B.BIN_DIR, "mglc_buoyancy")
    if not os.path.exists(exe):
        B.build_fortran_driver()
    return exe


def _initial(tmp_path, **nml):
    body = ", ".join(f"{k}={v}" for k, v in nml.items())
    path = tmp_path / "run.nml"
    path.write_text(f"&buoyancy\n {body}, output_dir='{tmp_path}' /\n")
    r = subprocess.run([_driver(), str(path), "--initial-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2:
            vals[parts[0]] = float(parts[1]) if parts[0] != "itc_max" else int(parts[1])
    n = nml["total_nx"] * nml["total_ny"] * nml["total_nz"]
    T = np.fromfile(tmp_path / "buoyancy-initial-T.bin", dtype=np.float64)
    assert T.size == n
    return vals, T.reshape(nml["total_nz"], nml["total_ny"], nml["total_nx"])


def test_splitmix_matches_published_sequence():
    """SplitMix64 (Steele, Lea, Flood 2014): from state 0 the first outputs are
    0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F; here the state
    after n increments is (index + seed) * golden, so index + seed = 1, 2, 3."""
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = lbm.splitmix_uniform((3,), seed=1, amp=1.0)
    for x, z in zip(u, want):
        assert x == 2.0 * ((z >> 11) * (1.0 / 9007199254740992.0)) - 1.0


def test_derived_parameters_match_module(tmp_path):
    """Reference defaults (module.F90:11-33): 51^3, Ra 1e6, Pr 0.71, Ma 0.1,
    Ekman 0.001; the Python configuration derives the same constants."""
    vals, _ = _initial(tmp_path, total_nx=51, total_ny=51, total_nz=51)
    d = lbm.ThermalConfig(51, 51, 51, Rayleigh=1e6, Prandtl=0.71, Mach=0.1, Ekman=0.001).derived()
    assert vals["tauf"] == d["tau"]
    assert vals["viscosity"] == d["viscosity"]
    assert vals["diffusivity"] == d["diffusivity"]
    assert vals["omegaRatating"] == d["omega_rot"]
    assert vals["paraA"] == d["para_a"]
    assert vals["gBeta"] == d["g_beta"]
    # tauf = 0.5 + 0.1 * 51 * sqrt(3 * 0.71 / 1e6) by hand
    assert abs(vals["tauf"] - (0.5 + 5.1 * (2.13e-6) ** 0.5)) < 1e-15
    assert vals["timeUnit"] == pytest.approx((51.0 / d["g_beta"]) ** 0.5, rel=1e-15)
    assert vals["itc_max"] == int(1000 * vals["timeUnit"])


def test_no_coriolis_when_ekman_not_positive(tmp_path):
    vals, _ = _initial(tmp_path, total_nx=9, total_ny=9, total_nz=9, Ekman=0.0)
    assert vals["omegaRatating"] == 0.0


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_initial_temperature_noise_matches_python(tmp_path, axis):
    nx, ny, nz, seed, amp = 7, 6, 5, 11, 1e-3
    _, T = _initial(tmp_path, total_nx=nx, total_ny=ny, total_nz=nz, heated_axis=axis, noise_amplitude=amp,
                    noise_seed=seed, Thot=0.5, Tcold=-0.5)
    if axis == 1:
        want = lbm.buoyancy_initial_temperature(nx, ny, nz, t_hot=0.5, t_cold=-0.5, seed=seed, amp=amp)
    else:
        want = np.zeros((nz, ny, nx))
        if axis == 0:
            want[:, :, 1:-1] = lbm.splitmix_uniform((nz, ny, nx - 2), seed, amp)
            want[:, :, 0], want[:, :, -1] = 0.5, -0.5
        else:
            want[1:-1] = lbm.splitmix_uniform((nz - 2, ny, nx), seed, amp)
            want[0], want[-1] = 0.5, -0.5
    assert np.array_equal(T, want)
    inner = np.delete(T, [0, T.shape[2 - axis] - 1], axis=2 - axis)
    assert np.all(np.abs(inner) < amp) and np.std(inner) > amp / 3


def test_initial_temperature_without_noise_is_held_layers(tmp_path):
    _, T = _initial(tmp_path, total_nx=5, total_ny=6, total_nz=4)
    want = np.zeros((4, 6, 5))
    want[:, 0, :], want[:, -1, :] = 1.0, 0.0
    assert np.array_equal(T, want)


def test_buoyancy_driver_binds_only_declared_symbols():
    from tests.test_capi import declared_functions
    out = subprocess.run(["nm", "-D", "--undefined-only", _driver()], capture_output=True, text=True,
                         check=True).stdout
    used = {l.split()[-1] for l in out.splitlines() if " lbm_" in l}
    assert {"lbm_check_thermal", "lbm_nusselt", "lbm_reynolds", "lbm_load_fields"} <= used
    assert used <= set(declared_functions())


def test_lbm_params_size_unchanged():
    assert ctypes.sizeof(lbm.LbmParams) == 152
