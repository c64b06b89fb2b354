"""CPU tests of the two-dimensional restatement of the reference's two-disc
program (MPI/Micro_particles/fortran/case4/seq/CaseIV_DKT.F90) against
states written by that program itself (tests/golden/dkt_s*.npz, made by
tests/golden/make_golden.py from the reference's own subroutines).

The restatement computes its per-link and per-particle formulas (calQ, the
interpolated bounce-back with the moving wall, the momentum exchange, the
springs, the refill extrapolation) with the functions of
oracle/particle_scheme.h that the three-dimensional particle oracle uses, so
bit-for-bit agreement here pins those functions for the 3-D oracle too.
"""
import hashlib
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEPS = [1, 300, 2400]


def _fixture(ns):
    return np.load(os.path.join(GOLDEN, f"dkt_s{ns}.npz"))


def _bits(x):
    return np.ascontiguousarray(x, dtype="<f8").view(np.uint64)


@pytest.fixture(scope="module")
def states(oracle_lib):
    return {ns: oracle_lib.dkt_run(ns, with_f=True) for ns in STEPS}


@pytest.mark.parametrize("ns", STEPS)
def test_dkt_disc_state_is_bit_exact(oracle_lib, states, ns):
    O = oracle_lib
    g = _fixture(ns)
    st = states[ns]
    scal = np.array([st["itc"], st["shift"], st["rho_avg"]] +
                    [x for name in O.DKT_SCALARS[3:] for x in st[name]], dtype=np.float64)
    assert int(g["steps"]) == ns and st["itc"] == ns
    assert np.array_equal(_bits(scal), _bits(g["scal"])), (scal, g["scal"])


@pytest.mark.parametrize("ns", STEPS)
def test_dkt_fields_are_bit_exact(states, ns):
    g = _fixture(ns)
    st = states[ns]
    for k in ("rho", "u", "v", "obst", "f"):
        got = hashlib.sha256(np.ascontiguousarray(st[k], dtype="<f8").tobytes()).hexdigest()
        assert got == str(g[f"sha_{k}"]), k
    # the windows at the discs, element by element (they name the first
    # differing node when a digest fails)
    for c in range(2):
        j0, i0 = (int(x) for x in g[f"origin{c}"])
        w = g["rho0"].shape[0]
        for k in ("rho", "u", "v"):
            assert np.array_equal(_bits(st[k][j0:j0 + w, i0:i0 + w]), _bits(g[f"{k}{c}"])), (k, c)


def test_dkt_fixtures_exercise_the_scheme():
    """The states cover the moving window and the particle spring: by step
    2400 the window has shifted and the discs are closer than
    r1 + r2 + threshold (10 + 10 + 4, CaseIV_DKT.F90:26, 754)."""
    g = _fixture(2400).get("scal")
    shift = int(g[1])
    xc, yc = g[3:5], g[5:7]
    dij = np.hypot(xc[0] - xc[1], yc[0] - yc[1])
    assert shift > 0
    assert 20.0 < dij < 24.0
    # both discs moved and spin: the moving-wall and rotation terms are live
    assert np.all(np.abs(g[9:11]) > 1e-3)      # vc
    assert np.all(np.abs(g[11:13]) > 0.0)      # omega


def test_dkt_scheme_mistake_is_seen(oracle_lib, states):
    """A perturbation of one population by one unit in the last place at
    step 300 reaches the fields and the digests tell it apart (the comparison
    is not vacuous)."""
    f = states[300]["f"].copy()
    f[400, 100, 3] = np.nextafter(f[400, 100, 3], np.inf)
    got = hashlib.sha256(np.ascontiguousarray(f, dtype="<f8").tobytes()).hexdigest()
    assert got != str(_fixture(300)["sha_f"])
