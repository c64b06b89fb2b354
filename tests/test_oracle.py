"""CPU tests of the oracle (the restated reference step), pinned against the
reference itself and against closed forms.  No GPU."""
import glob
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "cavity_mrt_*.npz")))


def relmax(a, b):
    """max |a - b| / max |b| -- the relative max error of the parity bar."""
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_fixture(oracle_lib, path):
    O = oracle_lib
    d = np.load(path)
    p = O.cavity_params(int(d["nx"]), int(d["ny"]), int(d["nz"]), float(d["tauf"]), float(d["U0"]),
                        rho0=float(d["rho0"]))
    rho, u, v, w, _ = O.cavity_run(p, int(d["steps"]), nthreads=1)
    for got, name in ((rho, "rho"), (u, "u"), (v, "v"), (w, "w")):
        assert relmax(got, d[name]) <= 1e-12, name


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 7


def test_basis_inverse_and_reference_coefficients(oracle_lib):
    M, Minv = oracle_lib.d3q19_basis()
    assert np.abs(M @ Minv - np.eye(19)).max() < 1e-15
    assert np.all(M == np.round(M))
    # inverse coefficients written out in collision.f90:118-189
    known = {(0, 0): 1 / 19, (0, 1): -5 / 399, (0, 2): 1 / 21, (1, 1): -11 / 2394, (1, 2): -1 / 63,
             (1, 3): 1 / 10, (1, 4): -1 / 10, (1, 9): 1 / 18, (1, 10): -1 / 18, (3, 11): 1 / 12,
             (3, 12): -1 / 12, (3, 9): -1 / 36, (7, 1): 4 / 1197, (7, 2): 1 / 252, (7, 4): 1 / 40,
             (7, 10): 1 / 72, (7, 12): 1 / 24, (7, 13): 1 / 4, (7, 16): 1 / 8, (7, 17): -1 / 8,
             (11, 15): 1 / 4, (11, 16): -1 / 8, (11, 18): 1 / 8, (15, 9): -1 / 18, (15, 10): -1 / 36,
             (15, 14): 1 / 4, (18, 17): -1 / 8, (18, 18): 1 / 8, (0, 3): 0.0, (5, 7): 1 / 10}
    for (a, k), val in known.items():
        assert Minv[a, k] == pytest.approx(val, abs=1e-17), (a, k)
    # rows 1 and 2 of the reference (collision.f90:24-30): -30, -11 x6, 8 x12 and 12, -4 x6, 1 x12
    assert list(M[1]) == [-30] + [-11] * 6 + [8] * 12
    assert list(M[2]) == [12] + [-4] * 6 + [1] * 12


def _feq(rho, u, v, w):
    ex = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
    ey = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
    ez = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
    wt = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    un = u * ex + v * ey + w * ez
    return rho * wt * (1 + 3 * un + 4.5 * un * un - 1.5 * (u * u + v * v + w * w))


def test_equilibrium_moments_match_reference_meq(oracle_lib):
    """M feq equals the reference's meq (collision.f90:73-91) with rho on row 12:
    two formulas of the reference (initial.f90:63-73 and collision.f90) agree."""
    M, _ = oracle_lib.d3q19_basis()
    rng = np.random.default_rng(7)
    for _ in range(20):
        rho = 1 + 0.1 * rng.standard_normal()
        u, v, w = 0.1 * rng.standard_normal(3)
        m = M @ _feq(rho, u, v, w)
        us2 = u * u + v * v + w * w
        meq = [rho, rho * (-11 + 19 * us2), rho * (3 - 5.5 * us2), rho * u, -2 / 3 * rho * u, rho * v,
               -2 / 3 * rho * v, rho * w, -2 / 3 * rho * w, rho * (2 * u * u - v * v - w * w),
               -0.5 * rho * (2 * u * u - v * v - w * w), rho * (v * v - w * w), -0.5 * rho * (v * v - w * w),
               rho * u * v, rho * v * w, rho * u * w, 0, 0, 0]
        np.testing.assert_allclose(m, meq, rtol=0, atol=1e-14)


def test_bgk_equals_mrt_with_uniform_rates(oracle_lib):
    """At tau = 1/2 + sqrt(3)/4, Sq = 8(2tau-1)/(8tau-1) equals Snu = 1/tau, so MRT
    with meq(12) carrying rho relaxes every non-conserved moment at 1/tau: that is
    BGK.  Pins the BGK branch (commented out in collision.f90:192-198) against
    the MRT branch, which the reference fixtures pin."""
    O = oracle_lib
    tau = 0.5 + np.sqrt(3) / 4
    assert abs(1 / tau - 8 * (2 * tau - 1) / (8 * tau - 1)) < 1e-15
    a = O.cavity_run(O.cavity_params(10, 9, 8, tau, 0.1, collision=O.MRT, meq12_rho=1), 60, nthreads=1)
    b = O.cavity_run(O.cavity_params(10, 9, 8, tau, 0.1, collision=O.BGK), 60, nthreads=1)
    for x, y in zip(a[:4], b[:4]):
        assert relmax(x, y) < 1e-12
    # and the cavity quirk (meq(12) without rho) does make a difference
    c = O.cavity_run(O.cavity_params(10, 9, 8, tau, 0.1, collision=O.MRT, meq12_rho=0), 60, nthreads=1)
    assert relmax(c[2], b[2]) > 1e-9


@pytest.mark.parametrize("coll", ["MRT", "BGK"])
def test_rest_state_is_fixed_point(oracle_lib, coll):
    O = oracle_lib
    p = O.cavity_params(6, 7, 5, 0.7, 0.0, collision=getattr(O, coll))
    rho, u, v, w, _ = O.cavity_run(p, 25, nthreads=1)
    assert np.abs(rho - 1).max() < 1e-15
    assert max(np.abs(x).max() for x in (u, v, w)) < 1e-16


@pytest.mark.parametrize("coll", ["MRT", "BGK"])
def test_mass_conservation(oracle_lib, coll):
    """Halfway bounce-back conserves mass and the lid terms -rho U0/6 and +rho U0/6
    cancel per node (bounce_back.f90:77-78)."""
    O = oracle_lib
    p = O.cavity_params(12, 11, 10, 0.6, 0.1, collision=getattr(O, coll))
    rho, *_ = O.cavity_run(p, 200, nthreads=1)
    assert abs(rho.sum() - rho.size) / rho.size < 1e-13


def test_mirror_symmetry_about_y_midplane(oracle_lib):
    """The lid moves along x, so the flow is mirror symmetric about y = ny/2:
    u, w even and v odd in y."""
    O = oracle_lib
    rho, u, v, w, _ = O.cavity_run(O.cavity_params(10, 12, 9, 0.6, 0.1), 150, nthreads=1)
    assert np.abs(u - u[:, ::-1, :]).max() < 1e-14
    assert np.abs(w - w[:, ::-1, :]).max() < 1e-14
    assert np.abs(v + v[:, ::-1, :]).max() < 1e-14
    assert np.abs(v).max() > 1e-4


def test_check_closed_form(oracle_lib):
    """check.f90:9-34 on hand-made fields: u = 2, up = 1 gives e1 = n, e2 = 4n."""
    import ctypes
    O = oracle_lib
    p = O.cavity_params(4, 3, 2, 0.6, 0.1)
    n = 24
    u = np.full(n, 2.0); v = np.zeros(n); w = np.full(n, 1e300 * 0)  # w = 0
    up = np.ones(n); vp = np.zeros(n)
    f = O.lib().oracle_cavity_check
    f.argtypes = [ctypes.POINTER(O.CavityParams)] + [ctypes.POINTER(ctypes.c_double)] * 5
    e = f(ctypes.byref(p), *(O._dp(a) for a in (u, v, w, up, vp)))
    assert e == pytest.approx(0.5, abs=1e-16)
    assert np.all(up == 2.0)
    e2 = f(ctypes.byref(p), *(O._dp(a) for a in (u, v, w, up, vp)))
    assert e2 == 0.0


# errorU at itc = 2000 of the reference program itself
# (MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked, N = 17, Re = 100, U0 = 0.1,
# so tauf = 3 U0 N / Re + 0.5 = 0.551; check.f90:9-34 after 2000 steps of
# main.f90:85-101), as printed by a one-rank run; runs on 2, 4 and 8 ranks
# printed 0.86857284647299815, ...915 and ...893 (the Allreduce order).
REFERENCE_ERROR_U_17_RE100_ITC2000 = 0.86857284647299848


def test_residual_matches_the_reference_programs_known_answer(oracle_lib):
    """The whole step chain and check over 2000 steps against the number the
    reference program prints, to the spread of its own rank counts."""
    O = oracle_lib
    N = 17
    p = O.cavity_params(N, N, N, 3 * 0.1 * N / 100 + 0.5, 0.1)
    errs = O.cavity_run(p, 2000, check_every=2000)[-1]
    assert len(errs) == 1
    assert errs[0] == pytest.approx(REFERENCE_ERROR_U_17_RE100_ITC2000, rel=1e-14)
