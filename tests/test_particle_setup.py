"""CPU tests of the host-side sphere placement used by the particle-laden
workload (seeded random sequential addition)."""
import numpy as np

from mglc_amd.lbm import random_spheres


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
