"""GPU parity tests of the D3Q19 cavity step, through the C-ABI.

FP64 storage must match the reference (golden fixtures made by the
reference's own Fortran) and the oracle within 1e-12 relative max error on
rho and u, v, w.  FP32/FP16 storage are checked against the FP64 GPU path
with the error bounds written in each test."""
import glob
import os

import numpy as np
import pytest

from mglc_amd.lbm import Cavity, CavityConfig, LbmError

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "cavity_mrt_*.npz")))
TOL = 1e-12  # FP64 parity bar: relative max error on rho and u


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def gpu_run(cfg, steps):
    print(f"[gpu] {cfg.collision} {cfg.storage} {cfg.nx}x{cfg.ny}x{cfg.nz} steps={steps}", flush=True)
    with Cavity(cfg) as c:
        c.initial()
        c.step(steps)
        return c.macro()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_fp64_mrt_matches_reference_fixture(path):
    d = np.load(path)
    cfg = CavityConfig(int(d["nx"]), int(d["ny"]), int(d["nz"]), U0=float(d["U0"]), tauf=float(d["tauf"]),
                       rho0=float(d["rho0"]), collision="mrt")
    got = gpu_run(cfg, int(d["steps"]))
    for g, name in zip(got, ("rho", "u", "v", "w")):
        assert relmax(g, d[name]) <= TOL, name


CASES = [
    # (nx, ny, nz, steps, tau, U0)
    (8, 9, 10, 1, 0.6, 0.1),
    (8, 9, 10, 2, 0.6, 0.1),
    (33, 20, 17, 200, 0.56, 0.1),
    (64, 48, 40, 300, 0.692, 0.1),
    (70, 65, 66, 100, 0.5195, 0.1),
]


@pytest.mark.parametrize("coll", ["mrt", "bgk"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[:3])) + f"_s{c[3]}")
def test_fp64_matches_oracle(oracle_lib, coll, case):
    nx, ny, nz, steps, tau, U0 = case
    O = oracle_lib
    p = O.cavity_params(nx, ny, nz, tau, U0, collision=O.MRT if coll == "mrt" else O.BGK)
    ref = O.cavity_run(p, steps, nthreads=0)[:4]
    got = gpu_run(CavityConfig(nx, ny, nz, U0=U0, tauf=tau, collision=coll), steps)
    for g, r, name in zip(got, ref, ("rho", "u", "v", "w")):
        assert relmax(g, r) <= TOL, name


def test_config_small_64cubed_bgk_5000_steps(oracle_lib):
    """The plain-method small case: 64^3, BGK FP64, U = 0.1, Re = 100,
    tau = 3 U N / Re + 1/2 = 0.692, 5000 steps, against the oracle."""
    O = oracle_lib
    tau = 3 * 0.1 * 64 / 100 + 0.5
    ref = O.cavity_run(O.cavity_params(64, 64, 64, tau, 0.1, collision=O.BGK), 5000, nthreads=0)[:4]
    got = gpu_run(CavityConfig(64, 64, 64, Reynolds=100.0, U0=0.1, collision="bgk"), 5000)
    for g, r, name in zip(got, ref, ("rho", "u", "v", "w")):
        assert relmax(g, r) <= TOL, name


def test_residual_sequence_matches_oracle(oracle_lib):
    """check.f90 every 100 steps: errorU from the GPU reduction equals the
    oracle's sequential sum."""
    O = oracle_lib
    nx, ny, nz, tau = 24, 20, 18, 0.6
    _, _, _, _, errs = O.cavity_run(O.cavity_params(nx, ny, nz, tau, 0.1), 500, check_every=100, nthreads=0)
    got = []
    with Cavity(CavityConfig(nx, ny, nz, tauf=tau)) as c:
        c.initial()
        for _ in range(5):
            c.step(100)
            got.append(c.check())
    np.testing.assert_allclose(got, errs, rtol=1e-10)


def test_residual_matches_the_reference_programs_known_answer():
    """errorU at itc = 2000 of the reference program itself
    (mpi_3d_blocked, 17^3, Re = 100, U0 = 0.1; check.f90:9-34), as a one-rank
    run prints it; its 2-, 4- and 8-rank runs differ from it by up to 4e-16
    relative."""
    with Cavity(CavityConfig(17, 17, 17, Reynolds=100.0, U0=0.1)) as c:
        c.initial()
        c.step(2000)
        e = c.check()
    assert e == pytest.approx(0.86857284647299848, rel=1e-12)


def test_mass_conserved_on_gpu():
    rho, *_ = gpu_run(CavityConfig(40, 36, 32, tauf=0.6), 400)
    assert abs(rho.sum() - rho.size) / rho.size < 1e-13


def test_nan_residual_is_reported():
    """A NaN in the fields makes check() fail with LBM_ERR_NAN (the reference
    silently leaves its loop, check.f90 / main.f90:85)."""
    n = (8, 8, 8)
    rho = np.ones(n); u = np.zeros(n); v = np.zeros(n); w = np.zeros(n)
    rho[3, 3, 3] = np.nan
    with Cavity(CavityConfig(8, 8, 8, tauf=0.6)) as c:
        c.load_fields(rho, u, v, w)
        c.step(5)
        with pytest.raises(LbmError) as ei:
            c.check()
        assert ei.value.status == 3


def test_load_fields_equals_initial():
    cfg = CavityConfig(16, 12, 10, tauf=0.6)
    n = (10, 12, 16)
    rho = np.ones(n); u = np.zeros(n); v = np.zeros(n); w = np.zeros(n)
    u[-1, :, :] = 0.1
    with Cavity(cfg) as c:
        c.load_fields(rho, u, v, w)
        c.step(30)
        a = c.macro()
    b = gpu_run(cfg, 30)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("coll", ["mrt", "bgk"])
@pytest.mark.parametrize("storage,l2_bound,mass_bound", [("f32", 1e-5, 1e-6), ("f16", 2e-3, 1e-3)])
def test_reduced_storage_against_fp64(coll, storage, l2_bound, mass_bound):
    """FP32 / FP16 storage of shifted populations (f - w), FP32 arithmetic,
    against the FP64 GPU path: relative L2 velocity error and mass drift."""
    base = dict(nx=48, ny=48, nz=48, Reynolds=400.0, U0=0.1, collision=coll)
    r64, u64, v64, w64 = gpu_run(CavityConfig(storage="f64", **base), 1000)
    r, u, v, w = gpu_run(CavityConfig(storage=storage, **base), 1000)
    vel64 = np.stack([u64, v64, w64])
    err = rel_l2(np.stack([u, v, w]), vel64)
    drift = abs(r.sum() - r64.sum()) / r64.sum()
    print(f"{storage} {coll}: L2 velocity error {err:.3e}, mass drift {drift:.3e}")
    assert err < l2_bound
    assert drift < mass_bound
