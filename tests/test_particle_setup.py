"""CPU tests of the host-side sphere placement: the particle-laden
workload's seeded random sequential addition, and the reference's jittered
grid (mpi_particle/initial.F90:48-84)."""
import numpy as np
import pytest

from mglc_amd.lbm import jittered_grid_spheres, random_spheres


def test_random_spheres_are_disjoint_inside_and_seeded():
    n = (64, 60, 70)
    xc, r = random_spheres(*n, count=120, rmin=2.0, rmax=4.0, seed=11, gap=1.0, min_gap=None)
    assert xc.shape == (120, 3) and r.shape == (120,)
    assert r.min() >= 2.0 and r.max() <= 4.0
    d = np.linalg.norm(xc[:, None, :] - xc[None, :, :], axis=2)
    need = r[:, None] + r[None, :] + 1.0
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= need)
    N = np.array(n, dtype=float)
    assert np.all(xc - r[:, None] - 1.0 >= 1.0)  # low walls at coordinate 1
    assert np.all(N - xc - r[:, None] >= 1.0)    # high walls at N
    xc2, r2 = random_spheres(*n, count=120, rmin=2.0, rmax=4.0, seed=11, gap=1.0, min_gap=None)
    assert np.array_equal(xc, xc2) and np.array_equal(r, r2)


def test_workload_packing_fraction_and_gaps():
    """192^3 with 2000 spheres of radius 4-8: about 29 % solid, relaxed to
    gaps of at least 2.5 lattice units, reproducibly."""
    xc, r = random_spheres(192, 192, 192, count=2000, rmin=4.0, rmax=8.0, seed=2025)
    frac = (4.0 / 3.0 * np.pi * r ** 3).sum() / 192 ** 3
    assert 0.25 < frac < 0.32
    from scipy.spatial import cKDTree
    pairs = cKDTree(xc).query_pairs(16.0 + 2.5, output_type="ndarray")
    gaps = np.linalg.norm(xc[pairs[:, 0]] - xc[pairs[:, 1]], axis=1) - r[pairs[:, 0]] - r[pairs[:, 1]]
    assert gaps.min() >= 2.5
    assert (xc - r[:, None] - 1.0).min() >= 2.5 and (192.0 - xc - r[:, None]).min() >= 2.5
    xc2, _ = random_spheres(192, 192, 192, count=2000, rmin=4.0, rmax=8.0, seed=2025)
    assert np.array_equal(xc, xc2)


def test_jittered_grid_follows_the_reference_layout():
    """Rows of 4 in a 201-wide box (x = 25, 75, 125, 175; 225 passes 200 and
    wraps, initial.F90:70-73), each centre within +-10 of its grid point,
    radius radius0 = 10, disjoint, reproducible for a seed."""
    xc, r = jittered_grid_spheres(201, 201, 101, count=10, radius=10.0, seed=3)
    grid = np.array([[25 + 50 * (q % 4), 25 + 50 * ((q // 4) % 4), 25] for q in range(10)], dtype=float)
    assert np.all(np.abs(xc - grid) <= 10.0)
    assert np.all(r == 10.0)
    d = np.linalg.norm(xc[:, None, :] - xc[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    assert d.min() >= 30.0 - 1e-12          # pitch 50 less two half-jitters
    again, _ = jittered_grid_spheres(201, 201, 101, count=10, radius=10.0, seed=3)
    assert np.array_equal(xc, again)
    other, _ = jittered_grid_spheres(201, 201, 101, count=10, radius=10.0, seed=4)
    assert not np.array_equal(xc, other)
    # a full x-y layer moves z
    xc16, _ = jittered_grid_spheres(201, 201, 101, count=17, seed=3)
    assert abs(xc16[16, 2] - 75.0) <= 10.0
    with pytest.raises(ValueError):
        jittered_grid_spheres(201, 201, 40, count=17)
