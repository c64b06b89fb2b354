"""CPU tests of the moving-sphere oracle (D3Q19 generalisation of the
reference's case4 particle coupling), pinned by brute force and closed forms.
Its per-link formulas are shared with the 2-D restatement of the reference's
CaseIV_DKT.F90, which reproduces that program bit for bit
(test_oracle_dkt.py); the 3-D assembly has no reference output."""
import numpy as np
import pytest

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])


def brute_map(n, xc, r):
    """Python floats are IEEE doubles with each operation rounded: the
    (x - c)**2 sums of initial.F90:78-89 at -O3, evaluated directly."""
    nx, ny, nz = n
    owner = -np.ones((nz, ny, nx), dtype=np.int32)
    for q, (c, rad) in enumerate(zip(xc, r)):
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    dx, dy, dz = (i + 1) - c[0], (j + 1) - c[1], (k + 1) - c[2]
                    if (dx * dx + dy * dy) + dz * dz <= rad * rad:
                        owner[k, j, i] = q
    return owner


def test_particle_map_is_bit_exact_against_brute_force(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(5)
    n = (14, 12, 13)
    xc = np.stack([rng.uniform(3, 11, 4), rng.uniform(3, 9, 4), rng.uniform(3, 10, 4)], axis=1)
    # include a centre exactly on a node and radii whose square is an integer (ties)
    xc[0] = (7.0, 6.0, 7.0)
    r = np.array([2.0, 1.5, np.sqrt(5.0), 2.25])
    p = O.particle_params(*n, 4, 0.6)
    got = O.particle_map(p, xc, r)
    assert np.array_equal(got, brute_map(n, xc, r))
    # lattice points within radius 2 of a node: 1 + 6 + 12 + 8 + 6
    alone = O.particle_map(O.particle_params(*n, 1, 0.6), xc[:1], r[:1])
    assert (alone == 0).sum() == 33


def test_calq_bisection_matches_closed_form(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(9)
    xc = np.array([[10.3, 9.7, 10.1]])
    r = np.array([4.6])
    checked = 0
    for _ in range(400):
        x = np.round(xc[0] + rng.uniform(-6, 6, 3))
        d2 = np.sum((x - xc[0]) ** 2)
        for a in range(1, 19):
            e = np.array([EX[a], EY[a], EZ[a]])
            inside_next = np.sum((x + e - xc[0]) ** 2) <= r[0] ** 2
            if d2 > r[0] ** 2 and inside_next:
                rc, q, x0, closed = O.calq(xc, r, 0, x, a)
                assert rc == 0 and 0.0 <= q <= 1.0
                assert abs(q - closed) < 1e-8
                assert abs(np.linalg.norm(x0 - xc[0]) - r[0]) < 1e-9
                checked += 1
    assert checked > 50


def _setup(oracle_lib, n=(20, 20, 22), tau=0.8, rho_solid=1.0, gravity=0.0):
    O = oracle_lib
    xc = np.array([[10.4, 10.6, 11.2]])
    r = np.array([4.3])
    return O, O.particle_params(*n, 1, tau, rho_solid=rho_solid, gravity=gravity), xc, r


def test_neutral_sphere_at_rest_stays_at_rest(oracle_lib):
    """Rest equilibrium is a fixed point of the interpolated bounce-back for any
    q, and the momentum exchange on a closed surface sums to zero."""
    O, p, xc, r = _setup(oracle_lib)
    out = O.particle_run(p, 30, xc, r)
    assert out["rc"] == 0
    assert np.abs(out["U"]).max() < 1e-13 and np.abs(out["Om"]).max() < 1e-13
    assert np.abs(out["F"]).max() < 1e-12
    fluid = out["owner"] < 0
    assert np.abs(out["u"][fluid]).max() < 1e-13
    assert np.abs(out["rho"][fluid] - 1.0).max() < 1e-13


def test_heavy_sphere_settles_and_drags_fluid(oracle_lib):
    O, p, xc, r = _setup(oracle_lib, rho_solid=1.5, gravity=1e-4)
    out = O.particle_run(p, 60, xc, r)
    assert out["rc"] == 0
    Uz = out["U"][0, 2]
    assert Uz < 0
    # net force below the free-fall weight (the fluid resists)
    V = 4 / 3 * np.pi * r[0] ** 3
    assert out["F"][0, 2] > -(1.5 - out["rho_avg"]) * V * 1e-4
    assert out["w"][out["owner"] < 0].min() < -1e-6
    assert np.abs(out["U"][0, :2]).max() < 1e-2 * abs(Uz)  # off-centre staircase drift only


def test_interpenetration_and_wall_penetration_stop_the_run(oracle_lib):
    O = oracle_lib
    p = O.particle_params(24, 24, 24, 2, 0.8)
    out = O.particle_run(p, 1, np.array([[8.0, 12.0, 12.0], [13.0, 12.0, 12.0]]), np.array([3.0, 3.0]))
    assert out["rc"] == -2
    p1 = O.particle_params(24, 24, 24, 1, 0.8)
    out = O.particle_run(p1, 1, np.array([[2.5, 12.0, 12.0]]), np.array([3.0]))
    assert out["rc"] == -3


def test_fields_are_conditioned_by_the_bisection_stop(oracle_lib):
    """calQ stops at |d - r| < 1e-9 (particle_bounceback.F90:109-116): a
    last-bit change of a centre can move q by ~2^-30 on some link, which
    moves the flow by far more than rounding.  Perturbing the centre by
    1e-15 relative changes the fields by more than 1e-13 but stays below
    1e-9 relative: the bar the GPU parity tests use for particle fields."""
    O = oracle_lib
    n = (24, 22, 30)
    xc, r = np.array([[12.3, 11.1, 17.6]]), np.array([4.7])
    p = O.particle_params(*n, 1, 0.7, gravity=2e-4)
    a = O.particle_run(p, 150, xc, r)
    b = O.particle_run(p, 150, xc * (1 + 1e-15), r)
    d = np.abs(a["u"] - b["u"]).max() / np.abs(a["u"]).max()
    assert 1e-13 < d < 1e-9
    assert np.abs(a["xc"] - b["xc"]).max() < 1e-12


def test_parallel_per_node_loops_change_no_value(oracle_lib):
    """The oracle's per-node loops (collision, streaming with the particle
    bounce-back, macro) run in parallel: one thread and many give the same
    bits in every field and every sphere quantity."""
    from mglc_amd.lbm import random_spheres
    O = oracle_lib
    xc, r = random_spheres(48, 44, 40, count=12, rmin=3.0, rmax=6.0, seed=7)
    p = O.particle_params(48, 44, 40, 12, 0.65, rho_solid=1.5, gravity=1e-4)
    O.set_threads(1)
    a = O.particle_run(p, 60, xc, r)
    O.set_threads(4)
    b = O.particle_run(p, 60, xc, r)
    import os
    O.set_threads(os.cpu_count() or 1)
    assert a["rc"] == 0 and b["rc"] == 0
    for k in ("xc", "U", "Om", "F", "T", "owner", "rho", "u", "v", "w"):
        assert np.array_equal(a[k], b[k]), k
