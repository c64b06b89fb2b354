"""CPU tests of the buoyancy-cavity oracle (D3Q19 + Guo forcing, D3Q7 MRT
temperature), pinned against the reference's own subroutines and closed forms."""
import glob
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "thermal_[0-9]*.npz")))
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def velocity_scale(d):
    """Buoyancy velocity sqrt(gBeta dT N) with the reference's N = 51 that
    gBeta was computed for (module.F90:11,28-29)."""
    return float(np.sqrt(d["gBeta"] * (d["Thot"] - d["Tcold"]) * 51))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_thermal_oracle_matches_reference_fixture(oracle_lib, path):
    """rho and T within 1e-12 relative max; velocities within 1e-12 of the
    buoyancy velocity scale (the flow starts at rest, so its own maximum is not
    a scale early on)."""
    O = oracle_lib
    d = np.load(path)
    p = O.thermal_params(int(d["nx"]), int(d["ny"]), int(d["nz"]), float(d["tauf"]), float(d["paraA"]),
                         float(d["gBeta"]), float(d["omegaRatating"]), float(d["Thot"]), float(d["Tcold"]),
                         float(d["Tref"]), tconst_axis=1)
    rho, u, v, w, T, _ = O.thermal_run(p, int(d["steps"]), nthreads=1)
    assert relmax(rho, d["rho"]) <= 1e-12
    assert relmax(T, d["T"]) <= 1e-12
    U = velocity_scale(d)
    for g, name in ((u, "u"), (v, "v"), (w, "w")):
        assert np.max(np.abs(g - d[name])) <= 1e-12 * U, name


def test_thermal_fixtures_present():
    assert len(GOLDEN) >= 5


def test_guo_source_is_moments_of_velocity_space_forcing(oracle_lib):
    """The moment-space source of evolution_f.F90:107-125 equals M applied to
    Guo's forcing term w_a [3 (e_a - u) + 9 (e_a . u) e_a] . F (Guo, Zheng & Shi,
    PRE 65, 046308, 2002)."""
    O = oracle_lib
    M, _ = O.d3q19_basis()
    rng = np.random.default_rng(11)
    for _ in range(20):
        u, v, w = 0.05 * rng.standard_normal(3)
        F = 1e-3 * rng.standard_normal(3)
        e = np.stack([EX, EY, EZ], axis=1).astype(float)
        U = np.array([u, v, w])
        Fa = W * ((3 * (e - U) + 9 * (e @ U)[:, None] * e) @ F)
        np.testing.assert_allclose(M @ Fa, O.guo_source_moments(u, v, w, *F), rtol=0, atol=1e-17)


def test_d3q7_basis_and_equilibrium(oracle_lib):
    O = oracle_lib
    N, Ninv = O.d3q7_basis()
    assert np.abs(N @ Ninv - np.eye(7)).max() < 1e-15
    # inverse coefficients written in evolution_g.F90:40-46
    assert Ninv[0, 0] == pytest.approx(1 / 7) and Ninv[0, 4] == pytest.approx(-1 / 7)
    assert Ninv[1, 1] == pytest.approx(0.5) and Ninv[1, 4] == pytest.approx(1 / 42)
    assert Ninv[1, 5] == pytest.approx(1 / 6) and Ninv[3, 5] == pytest.approx(-1 / 12)
    assert Ninv[3, 6] == pytest.approx(0.25) and Ninv[5, 6] == pytest.approx(-0.25)
    # N applied to the initial g of initial.F90:176-179 gives neq of evolution_g.F90:23-29
    rng = np.random.default_rng(5)
    for _ in range(10):
        a = -6 + 8 * rng.random()
        T = rng.standard_normal()
        U = 0.05 * rng.standard_normal(3)
        wt = np.array([(1 - a) / 7] + [(a + 6) / 42] * 6)
        un = EX[:7] * U[0] + EY[:7] * U[1] + EZ[:7] * U[2]
        geq = wt * T * (1 + 21 / (6 + a) * un)
        np.testing.assert_allclose(N @ geq, [T, T * U[0], T * U[1], T * U[2], T * a, 0, 0], atol=1e-15)


def test_conduction_steady_state_is_linear_with_unit_nusselt(oracle_lib):
    """Without buoyancy the temperature between walls held at Thot (y = 0) and
    Tcold (y = ny) relaxes to the linear conduction profile, and both Nusselt
    definitions give 1."""
    O = oracle_lib
    nx, ny, nz, kappa = 4, 16, 4, 0.1
    para_a = 42 * np.sqrt(3) * kappa - 6
    p = O.thermal_params(nx, ny, nz, 0.8, para_a, 0.0, 0.0, 0.5, -0.5, 0.0, tconst_axis=1)
    rho, u, v, w, T, _ = O.thermal_run(p, 12000, nthreads=1)
    y = np.arange(ny) + 0.5
    expect = 0.5 - y / ny
    assert np.abs(T - expect[None, :, None]).max() < 1e-10
    nu_vol, nu_wall = O.thermal_nusselt(p, w, T, kappa)
    assert nu_vol == pytest.approx(1.0, abs=1e-12)
    assert nu_wall == pytest.approx(1.0, abs=1e-9)


def test_thermal_residuals_and_coriolis(oracle_lib):
    """check.F90 residuals are produced every check step, and the Coriolis term
    changes the flow only when its rate is non-zero."""
    O = oracle_lib
    d = O.thermal_derived(16, 1e5, 0.71, 0.1, Ekman=0.001)
    base = dict(tau=d["tau"], para_a=d["para_a"], g_beta=d["g_beta"])
    a = O.thermal_run(O.thermal_params(12, 16, 14, omega_rot=0.0, **base), 400, check_every=100, nthreads=1)
    b = O.thermal_run(O.thermal_params(12, 16, 14, omega_rot=d["omega_rot"], **base), 400, check_every=100,
                      nthreads=1)
    assert a[5].shape == (4, 2) and np.all(a[5] > 0)
    assert relmax(b[1], a[1]) > 1e-6
