"""GPU parity tests of the buoyancy-driven cavity (D3Q19 MRT + Guo force,
D3Q7 MRT temperature) through the C-ABI: against fixtures written by the
reference's own subroutines, against the oracle at the buoyancy-cavity
parameters (Ra = 1e5, Pr = 0.71, walls at +-0.5) with a seeded noisy start,
and against closed forms.

Parity bar (FP64): rho and T within 1e-12 relative max; velocities within
1e-12 of the buoyancy velocity scale sqrt(gBeta dT N) (the flow starts at
rest, so the instantaneous maximum is not a scale early on)."""
import glob
import os

import numpy as np
import pytest

from mglc_amd.lbm import Cavity, LbmParams, ThermalConfig, buoyancy_initial_temperature

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "thermal_[0-9]*.npz")))


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def params_from_fixture(d):
    p = LbmParams()
    p.nx, p.ny, p.nz = int(d["nx"]), int(d["ny"]), int(d["nz"])
    p.thermal, p.thermal_axis = 1, 1
    p.tau, p.u_lid, p.rho0 = float(d["tauf"]), 0.0, 1.0
    p.para_a, p.g_beta, p.omega_rot = float(d["paraA"]), float(d["gBeta"]), float(d["omegaRatating"])
    p.t_hot, p.t_cold, p.t_ref = float(d["Thot"]), float(d["Tcold"]), float(d["Tref"])
    return p


def gpu_thermal(p_or_cfg, steps, fields=None):
    with Cavity(p_or_cfg) as c:
        if fields is None:
            c.initial()
        else:
            c.load_fields(*fields)
        c.step(steps)
        return (*c.macro(), c.temperature())


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_fp64_thermal_matches_reference_fixture(path):
    """Reference parameters: Ra = 1e6, Pr = 0.71, Ekman = 0.001 (Coriolis on)."""
    d = np.load(path)
    print(f"[gpu] thermal fixture {os.path.basename(path)}", flush=True)
    rho, u, v, w, T = gpu_thermal(params_from_fixture(d), int(d["steps"]))
    U = float(np.sqrt(d["gBeta"] * (d["Thot"] - d["Tcold"]) * 51))
    assert relmax(rho, d["rho"]) <= 1e-12
    assert relmax(T, d["T"]) <= 1e-12
    for g, name in ((u, "u"), (v, "v"), (w, "w")):
        assert np.max(np.abs(g - d[name])) <= 1e-12 * U, name


def c2_like(nx, ny, nz, **kw):
    """Side-heated cavity of the buoyancy case: Ra = 1e5, Pr = 0.71, Ma = 0.1,
    y walls at +0.5 / -0.5, other walls adiabatic, all no-slip, no rotation."""
    return ThermalConfig(nx, ny, nz, Rayleigh=1e5, Prandtl=0.71, Thot=0.5, Tcold=-0.5, thermal_axis=1, **kw)


@pytest.mark.parametrize("shape,steps", [((24, 20, 18), 300), ((33, 32, 30), 800)])
def test_fp64_thermal_matches_oracle_with_noise(oracle_lib, shape, steps):
    O = oracle_lib
    nx, ny, nz = shape
    cfg = c2_like(nx, ny, nz)
    d = cfg.derived()
    T0 = buoyancy_initial_temperature(nx, ny, nz, seed=2024)
    z = np.zeros_like(T0)
    fields = (np.ones_like(T0), z, z, z, T0)
    p = O.thermal_params(nx, ny, nz, d["tau"], d["para_a"], d["g_beta"], 0.0, 0.5, -0.5, 0.0, tconst_axis=1)
    ref = O.thermal_run(p, steps, fields=fields, check_every=steps // 2)
    print(f"[gpu] thermal oracle {shape} steps={steps}", flush=True)
    with Cavity(cfg) as c:
        c.load_fields(*fields)
        c.step(steps // 2)
        e1 = c.check_thermal()
        c.step(steps - steps // 2)
        e2 = c.check_thermal()
        got = (*c.macro(), c.temperature())
        nu = c.nusselt()
    U = float(np.sqrt(d["g_beta"] * 1.0 * nz))
    assert relmax(got[0], ref[0]) <= 1e-12
    assert relmax(got[4], ref[4]) <= 1e-12
    for g, r, name in zip(got[1:4], ref[1:4], "uvw"):
        assert np.max(np.abs(g - r)) <= 1e-12 * U, name
    np.testing.assert_allclose([e1, e2], ref[5], rtol=1e-9)
    ref_nu = O.thermal_nusselt(p, ref[3], ref[4], d["diffusivity"])
    np.testing.assert_allclose(nu, ref_nu, rtol=1e-11)


def test_conduction_closed_form_on_gpu():
    """No buoyancy: linear conduction profile between the held walls, Nu = 1."""
    nx, ny, nz, kappa = 8, 16, 8, 0.1
    p = ThermalConfig(nx, ny, nz, thermal_axis=1).params()
    p.g_beta = 0.0
    p.para_a = 42 * np.sqrt(3) * kappa - 6
    p.tau = 0.8
    rho, u, v, w, T = gpu_thermal(p, 12000)
    y = np.arange(ny) + 0.5
    assert np.abs(T - (0.5 - y / ny)[None, :, None]).max() < 1e-10
    with Cavity(p) as c:
        c.initial()
        c.step(12000)
        nu_vol, nu_wall = c.nusselt()
    assert nu_vol == pytest.approx(1.0, abs=1e-12)
    assert nu_wall == pytest.approx(1.0, abs=1e-9)


def test_side_heated_cavity_is_centro_symmetric():
    """Hot and cold walls at +-0.5 with Tref = 0: the flow is symmetric under the
    point reflection through the cavity centre across the y-z plane
    (y -> ny-1-y, z -> nz-1-z) with T, v, w changing sign."""
    rho, u, v, w, T = gpu_thermal(c2_like(16, 24, 24), 3000)
    Tr = T[::-1, ::-1, :]
    assert np.abs(T + Tr).max() < 1e-12
    assert np.abs(v + v[::-1, ::-1, :]).max() < 1e-13
    assert np.abs(w + w[::-1, ::-1, :]).max() < 1e-13
    assert np.abs(w).max() > 1e-5


def test_thermal_mass_conserved():
    rho, *_ = gpu_thermal(c2_like(20, 24, 22), 1500)
    assert abs(rho.sum() - rho.size) / rho.size < 1e-13


def test_full_size_buoyancy_cavity_against_oracle_fixture():
    """The buoyancy-cavity workload itself: 128^3, Ra = 1e5, Pr = 0.71, walls
    at +-0.5, seeded noise, 20000 steps, against the oracle's run of the same
    case (tests/golden/thermal_full_128_s20000.npz: Nusselt numbers, residual
    history, mid-line profiles and a strided subsample of the fields)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "thermal_full_128_s20000.npz")
    d = np.load(path)
    n, steps, stride = int(d["N"]), int(d["steps"]), int(d["stride"])
    cfg = c2_like(n, n, n)
    der = cfg.derived()
    T0 = buoyancy_initial_temperature(n, n, n, seed=int(d["seed"]))
    z = np.zeros_like(T0)
    print(f"[gpu] thermal full {n}^3 steps={steps}", flush=True)
    errs = []
    with Cavity(cfg) as c:
        c.load_fields(np.ones_like(T0), z, z, z, T0)
        for _ in range(steps // 2000):
            c.step(2000)
            errs.append(c.check_thermal())
        rho, u, v, w = c.macro()
        T = c.temperature()
        nu = c.nusselt()
    U = float(np.sqrt(der["g_beta"] * 1.0 * n))
    s = slice(stride // 2, None, stride)
    h = n // 2
    np.testing.assert_allclose(nu, (float(d["nu_vol"]), float(d["nu_wall"])), rtol=1e-11)
    np.testing.assert_allclose(np.array(errs), d["errs"], rtol=1e-8)
    assert relmax(rho[s, s, s], d["rho_sub"]) <= 1e-12
    assert relmax(T[s, s, s], d["T_sub"]) <= 1e-11
    for g, name in ((u, "u_sub"), (v, "v_sub"), (w, "w_sub")):
        assert np.max(np.abs(g[s, s, s] - d[name])) <= 1e-11 * U, name
    assert np.max(np.abs(T[h, :, h] - d["T_mid"])) <= 1e-11
    assert abs(float(rho.sum()) - float(d["mass"])) / float(d["mass"]) <= 1e-13


def test_reynolds_number_against_the_oracle_fields(oracle_lib):
    """lbm_reynolds = sqrt(<|u|^2>) nz / nu (RaNu.F90:36-44), from the GPU's
    FP64 sums, against the same formula on the oracle's fields."""
    O = oracle_lib
    nx, ny, nz, steps = 22, 20, 18, 600
    cfg = c2_like(nx, ny, nz)
    d = cfg.derived()
    T0 = buoyancy_initial_temperature(nx, ny, nz, seed=2024)
    z = np.zeros_like(T0)
    fields = (np.ones_like(T0), z, z, z, T0)
    p = O.thermal_params(nx, ny, nz, d["tau"], d["para_a"], d["g_beta"], 0.0, 0.5, -0.5, 0.0, tconst_axis=1)
    ref = O.thermal_run(p, steps, fields=fields)
    print(f"[gpu] reynolds {nx}x{ny}x{nz} steps={steps}", flush=True)
    with Cavity(cfg) as c:
        c.load_fields(*fields)
        c.step(steps)
        re = c.reynolds()
    u, v, w = ref[1], ref[2], ref[3]
    expect = float(np.sqrt(np.mean(u * u + v * v + w * w)) * nz / d["viscosity"])
    assert re > 0 and abs(re - expect) / expect <= 1e-10
