"""GPU parity tests of the moving-sphere coupling (D3Q19 generalisation of
the reference's MPI/Micro_particles case4 scheme) through the C-ABI.

Bars: the particle-to-node map is bit-exact (against the oracle's map and
against brute-force enumeration of the same centres); sphere positions,
velocities and angular velocities within 1e-12 relative of the oracle (force
and torque within 1e-10 of the largest single-link momentum exchange, since
they are sums of cancelling terms); FP64 fields within FIELD_TOL = 1e-9
relative max of the oracle.  The field bar is set by the reference's wall-
distance bisection, not by arithmetic: calQ stops at |d - r| < 1e-9
(particle_bounceback.F90:109-116), so centres that differ in the last bit
(FMA contraction, the order of the momentum-exchange sums) can stop it one
halving apart, moving q by ~2^-30; the oracle itself moves its fields by
~5e-11 relative when a centre is perturbed by 1e-15 (tested on the CPU in
test_oracle_particles.py).  FP32 storage is checked against the FP64 GPU path
with the bound in the test.  The reference's particle code is 2-D: the
oracle's per-link formulas are those its 2-D restatement of CaseIV_DKT.F90
reproduces bit for bit (test_oracle_dkt.py); the 3-D assembly around them
has no reference output and is pinned by the oracle's closed-form checks."""
import numpy as np
import pytest

from mglc_amd.lbm import Cavity, CavityConfig, LbmError, jittered_grid_spheres, random_spheres

pytestmark = pytest.mark.gpu

LBM_ERR_ARG, LBM_ERR_STATE, LBM_ERR_UNSUPPORTED, LBM_ERR_PARTICLE = 1, 6, 7, 8
FIELD_TOL = 1e-9


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def box(nx, ny, nz, tau=0.8, storage="f64", collision="mrt"):
    return CavityConfig(nx, ny, nz, U0=0.0, tauf=tau, meq12_rho=1, storage=storage, collision=collision)


def gpu_particles(cfg, xc, r, steps, U=None, Om=None, per_call=None, **kw):
    print(f"[gpu] particles {cfg.storage} {cfg.nx}x{cfg.ny}x{cfg.nz} np={len(r)} steps={steps}", flush=True)
    with Cavity(cfg) as c:
        c.particles_init(xc, r, U=U, omega=Om, **kw)
        owner0 = c.particle_map()
        per_call = per_call or steps
        for q in range(0, steps, per_call):
            c.step(min(per_call, steps - q))
        out = c.particles()
        out["owner0"] = owner0
        out["owner"] = c.particle_map()
        out["rho"], out["u"], out["v"], out["w"] = c.macro()
    return out


def brute_map(shape, xc, r):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    owner = -np.ones(shape, dtype=np.int32)
    for q, (c, rad) in enumerate(zip(xc, r)):
        dx, dy, dz = (i + 1.0) - c[0], (j + 1.0) - c[1], (k + 1.0) - c[2]
        owner[(dx * dx + dy * dy) + dz * dz <= rad * rad] = q
    return owner


THREE = dict(
    n=(34, 30, 36),
    xc=np.array([[10.3, 9.6, 12.2], [22.7, 18.4, 14.9], [15.2, 20.1, 25.4]]),
    r=np.array([4.3, 3.6, 5.1]),
)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_initial_map_is_bit_exact(oracle_lib, storage):
    O = oracle_lib
    n, xc, r = THREE["n"], THREE["xc"], THREE["r"]
    with Cavity(box(*n, storage=storage)) as c:
        c.particles_init(xc, r)
        got = c.particle_map()
        rho, u, v, w = c.macro()
    ref = O.particle_map(O.particle_params(*n, len(r), 0.8), xc, r)
    assert np.array_equal(got, ref)
    assert np.array_equal(got, brute_map(got.shape, xc, r))
    assert np.all(rho[got >= 0] == 1.5) and np.all(rho[got < 0] == 1.0)
    assert not np.any(u) and not np.any(v) and not np.any(w)


def test_fp64_moving_spheres_match_oracle(oracle_lib):
    """Three spheres launched with linear and angular velocity under gravity:
    interpolated bounce-back with the moving-wall term on both q branches,
    momentum exchange, springs, kinematics, map rebuild and many refills."""
    O = oracle_lib
    n, xc, r = THREE["n"], THREE["xc"], THREE["r"]
    U = np.array([[0.02, 0.011, -0.015], [-0.018, 0.004, 0.012], [0.006, -0.02, -0.01]])
    Om = np.array([[0.0, 0.0, 0.004], [0.003, -0.002, 0.0], [0.0, 0.005, 0.001]])
    steps, g = 120, 1e-4
    got = gpu_particles(box(*n), xc, r, steps, U=U, Om=Om, gravity=g)
    ref = O.particle_run(O.particle_params(*n, 3, 0.8, gravity=g), steps, xc, r, U=U, Om=Om)
    assert ref["rc"] == 0
    # the spheres moved across nodes: the map changed and refills happened
    assert np.count_nonzero(got["owner0"] != got["owner"]) > 100
    assert np.array_equal(got["owner"], ref["owner"])
    assert np.array_equal(got["owner"], brute_map(got["owner"].shape, got["xc"], r))
    for name in ("rho", "u", "v", "w"):
        assert relmax(got[name], ref[name]) <= FIELD_TOL, name
    for name in ("xc", "U", "Om"):
        assert relmax(got[name], ref[name]) <= 1e-12, name
    link_scale = 1.0 / 18.0  # largest f in a link term
    for name in ("F", "T"):
        assert np.abs(got[name] - ref[name]).max() <= 1e-10 * link_scale * r.max(), name


def test_fast_spheres_match_oracle(oracle_lib):
    """Two spheres launched at 0.30 and 0.26 nodes per step: a sphere that
    moves a quarter node or more in a step makes the map update rebuild every
    sphere's link list for the next step (the lists are otherwise kept while
    the map near their sphere is unchanged).  24 steps: the first eager, a
    graph replay of ten, the rest direct."""
    O = oracle_lib
    n, xc, r = THREE["n"], THREE["xc"], THREE["r"]
    U = np.array([[0.30, 0.0, 0.0], [0.0, 0.0, 0.0], [-0.26, 0.0, 0.0]])
    steps, g = 24, 1e-4
    got = gpu_particles(box(*n), xc, r, steps, U=U, gravity=g)
    ref = O.particle_run(O.particle_params(*n, 3, 0.8, gravity=g), steps, xc, r, U=U)
    assert ref["rc"] == 0
    assert np.abs(got["xc"][0, 0] - xc[0, 0]) > 2.5  # it did move fast
    assert np.array_equal(got["owner"], ref["owner"])
    assert np.array_equal(got["owner"], brute_map(got["owner"].shape, got["xc"], r))
    for name in ("rho", "u", "v", "w"):
        assert relmax(got[name], ref[name]) <= FIELD_TOL, name
    for name in ("xc", "U", "Om"):
        assert relmax(got[name], ref[name]) <= 1e-12, name
    link_scale = 1.0 / 18.0
    for name in ("F", "T"):
        assert np.abs(got[name] - ref[name]).max() <= 1e-10 * link_scale * r.max(), name


def test_fp64_jittered_grid_with_springs_matches_oracle(oracle_lib):
    """Eight spheres on the reference's jittered grid layout
    (mpi_particle/initial.F90:48-84), packed so that particle-particle gaps
    (2.3) and wall gaps (2.2) start inside the springs' threshold of 4:
    both spring forces act from the first step."""
    O = oracle_lib
    n = (62, 36, 24)
    xc, r = jittered_grid_spheres(*n, count=8, radius=5.0, pitch=14.0, jitter=4.0, seed=5)
    xc = xc + 3.0
    d = np.linalg.norm(xc[:, None] - xc[None], axis=2)
    np.fill_diagonal(d, np.inf)
    assert d.min() - 2 * 5.0 < 4.0 and (xc - r[:, None] - 1.0).min() < 4.0
    steps, g = 80, 1e-4
    print(f"[gpu] particles jittered grid n={n} steps={steps}", flush=True)
    got = gpu_particles(box(*n), xc, r, steps, gravity=g)
    ref = O.particle_run(O.particle_params(*n, 8, 0.8, gravity=g), steps, xc, r)
    assert ref["rc"] == 0
    assert np.array_equal(got["owner"], ref["owner"])
    for name in ("rho", "u", "v", "w"):
        assert relmax(got[name], ref[name]) <= FIELD_TOL, name
    for name in ("xc", "U", "Om"):
        assert relmax(got[name], ref[name]) <= 1e-12, name


@pytest.mark.parametrize("collision", ["mrt", "bgk"])
def test_fp64_settling_sphere_matches_oracle(oracle_lib, collision):
    O = oracle_lib
    n = (24, 22, 30)
    xc, r = np.array([[12.3, 11.1, 17.6]]), np.array([4.7])
    steps, g = 150, 2e-4
    got = gpu_particles(box(*n, tau=0.7, collision=collision), xc, r, steps, gravity=g)
    ref = O.particle_run(O.particle_params(*n, 1, 0.7, gravity=g, collision=O.BGK if collision == "bgk" else O.MRT),
                         steps, xc, r)
    assert ref["rc"] == 0 and got["U"][0, 2] < 0
    assert np.array_equal(got["owner"], ref["owner"])
    assert relmax(got["xc"], ref["xc"]) <= 1e-12
    for name in ("rho", "u", "v", "w"):
        assert relmax(got[name], ref[name]) <= FIELD_TOL, name
    # the sphere velocity against the settling velocity scale Us = sqrt(g r (rho_s - 1))
    # and the (symmetry-breaking, ~1e-7) angular velocity against Us / r
    Us = float(np.sqrt(g * r[0] * 0.5))
    assert np.abs(got["U"] - ref["U"]).max() <= 1e-12 * Us
    assert np.abs(got["Om"] - ref["Om"]).max() <= 1e-12 * Us / r[0]


def test_neutral_spheres_stay_at_rest():
    """rho_s = rho0: rest equilibrium is a fixed point of the interpolated
    bounce-back for every q, and the momentum exchange over a closed surface
    cancels."""
    n, xc, r = THREE["n"], THREE["xc"], THREE["r"]
    got = gpu_particles(box(*n), xc, r, 50, rho_solid=1.0)
    assert np.abs(got["U"]).max() < 1e-14 and np.abs(got["Om"]).max() < 1e-14
    fluid = got["owner"] < 0
    assert np.abs(got["u"][fluid]).max() < 1e-14
    assert np.abs(got["rho"][fluid] - 1.0).max() < 1e-13
    assert np.array_equal(got["owner"], got["owner0"])


def test_runs_are_deterministic():
    n, xc, r = THREE["n"], THREE["xc"], THREE["r"]
    a = gpu_particles(box(*n), xc, r, 40, gravity=1e-4)
    b = gpu_particles(box(*n), xc, r, 40, gravity=1e-4)
    for name in ("xc", "U", "Om", "F", "T", "rho", "u"):
        assert np.array_equal(a[name], b[name]), name


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_graph_replays_equal_direct_launches(storage):
    """A step call replays graphs of ten steps for the steps before its last;
    one step per call launches every kernel directly.  Both give the same
    state bit for bit (moving spheres: links, refills and map changes inside
    the replays)."""
    n, xc, r = THREE["n"], THREE["xc"], THREE["r"]
    U = np.array([[0.02, 0.011, -0.015], [-0.018, 0.004, 0.012], [0.006, -0.02, -0.01]])
    Om = np.array([[0.0, 0.0, 0.004], [0.003, -0.002, 0.0], [0.0, 0.005, 0.001]])
    a = gpu_particles(box(*n, storage=storage), xc, r, 37, U=U, Om=Om, gravity=1e-4)
    b = gpu_particles(box(*n, storage=storage), xc, r, 37, U=U, Om=Om, gravity=1e-4, per_call=1)
    c = gpu_particles(box(*n, storage=storage), xc, r, 37, U=U, Om=Om, gravity=1e-4, per_call=12)
    assert np.count_nonzero(a["owner0"] != a["owner"]) > 20
    for name in ("xc", "U", "Om", "F", "T", "owner", "rho", "u", "v", "w"):
        assert np.array_equal(a[name], b[name]), name
        assert np.array_equal(a[name], c[name]), name


def test_stop_step_inside_a_graph_replay():
    """Two spheres 1 node apart closing at 0.25 nodes per step interpenetrate
    at a step inside the first graph replay (steps 1-10 of a 40-step call):
    the reported step is the one direct launches report."""
    xc = np.array([[8.0, 12.0, 12.0], [15.0, 12.0, 12.0]])
    r = np.array([3.0, 3.0])
    U = np.array([[0.125, 0.0, 0.0], [-0.125, 0.0, 0.0]])
    msgs = []
    for per_call in (1, 40):
        with Cavity(box(24, 24, 24)) as c:
            c.particles_init(xc, r, U=U)
            with pytest.raises(LbmError) as e:
                for _ in range(0, 40, per_call):
                    c.step(per_call)
            assert e.value.status == LBM_ERR_PARTICLE and "interpenetration" in str(e.value)
            msgs.append(str(e.value))
    step = int(msgs[0].split("step ")[1].split(":")[0])
    assert 2 <= step <= 10, msgs[0]  # not the first step, not after the replay
    assert msgs[0] == msgs[1]


def test_fp32_trajectories_against_fp64():
    n = (48, 44, 52)
    xc, r = random_spheres(*n, count=12, rmin=3.0, rmax=5.0, seed=7, gap=3.0)
    steps, g = 600, 1e-4
    d = gpu_particles(box(*n), xc, r, steps, gravity=g)
    s = gpu_particles(box(*n, storage="f32"), xc, r, steps, gravity=g)
    disp = np.linalg.norm(d["xc"] - xc, axis=1)
    err = np.linalg.norm(s["xc"] - d["xc"], axis=1)
    print(f"[gpu] fp32 trajectory: max displacement {disp.max():.4g}, max deviation {err.max():.3g}", flush=True)
    assert disp.max() > 0.05
    assert err.max() <= 1e-3 * disp.max()
    assert np.abs(s["U"] - d["U"]).max() <= 1e-3 * np.abs(d["U"]).max()
    assert np.array_equal(s["owner"], brute_map(s["owner"].shape, s["xc"], r))


def test_stop_codes():
    # interpenetration (particle_force.F90:117-124), reported with its step;
    # the context then refuses to step until the spheres are attached again
    with Cavity(box(24, 24, 24)) as c:
        c.particles_init(np.array([[8.0, 12.0, 12.0], [13.0, 12.0, 12.0]]), np.array([3.0, 3.0]))
        with pytest.raises(LbmError) as e:
            c.step(3)
        assert e.value.status == LBM_ERR_PARTICLE and "step 1:" in str(e.value)
        assert "interpenetration" in str(e.value)
        with pytest.raises(LbmError) as e:
            c.step(1)
        assert e.value.status == LBM_ERR_STATE
        c.particles_init(np.array([[8.0, 12.0, 12.0], [16.0, 12.0, 12.0]]), np.array([3.0, 3.0]))
        c.step(2)
    # wall penetration (particle_force.F90:136-186)
    with Cavity(box(24, 24, 24)) as c:
        c.particles_init(np.array([[2.5, 12.0, 12.0]]), np.array([3.0]))
        with pytest.raises(LbmError) as e:
            c.step(1)
        assert e.value.status == LBM_ERR_PARTICLE and "wall" in str(e.value)


def test_bad_arguments_and_unsupported_contexts():
    with Cavity(box(16, 16, 16)) as c:
        with pytest.raises(LbmError) as e:
            c.particles_init(np.array([[8.0, 8.0, 8.0]]), np.array([-1.0]))
        assert e.value.status == LBM_ERR_ARG
    with Cavity(box(16, 16, 16, storage="f16")) as c:
        with pytest.raises(LbmError) as e:
            c.particles_init(np.array([[8.0, 8.0, 8.0]]), np.array([3.0]))
        assert e.value.status == LBM_ERR_UNSUPPORTED
    with Cavity(CavityConfig(16, 16, 16)) as c:  # moving lid
        with pytest.raises(LbmError) as e:
            c.particles_init(np.array([[8.0, 8.0, 8.0]]), np.array([3.0]))
        assert e.value.status == LBM_ERR_UNSUPPORTED


def test_workload_at_full_size(oracle_lib):
    """The particle-laden workload as specified: 192^3, 2000 spheres r in
    [4, 8] (seed 2025), rho_s / rho_f = 1.5, g = 1e-5 along -z, 5000 steps,
    FP32 storage against the FP64 path from the same start.  The maps of both
    runs equal the oracle's map of each run's own centres bit for bit, and the
    FP32 trajectories stay within 2e-5 of the FP64 displacement (measured
    4.5e-6) and the velocities within 1e-4 of the largest FP64 speed
    (measured 1.1e-5)."""
    O = oracle_lib
    n = (192, 192, 192)
    xc, r = random_spheres(*n, count=2000, rmin=4.0, rmax=8.0, seed=2025)
    P = O.particle_params(*n, 2000, 0.65)
    out = {}
    for storage in ("f32", "f64"):
        print(f"[gpu] particle workload 192^3, 2000 spheres, {storage}, 5000 steps", flush=True)
        with Cavity(box(*n, tau=0.65, storage=storage)) as c:
            c.particles_init(xc, r, rho_solid=1.5, gravity=1e-5)
            assert np.array_equal(c.particle_map(), O.particle_map(P, xc, r))
            c.step(5000)
            st = c.particles()
            owner = c.particle_map()
        assert np.all(np.isfinite(st["xc"])) and np.all(np.isfinite(st["U"]))
        assert np.array_equal(owner, O.particle_map(P, st["xc"], r)), storage
        out[storage] = st
    disp = np.linalg.norm(out["f64"]["xc"] - xc, axis=1).max()
    dev = np.linalg.norm(out["f32"]["xc"] - out["f64"]["xc"], axis=1).max()
    dU = np.abs(out["f32"]["U"] - out["f64"]["U"]).max() / np.abs(out["f64"]["U"]).max()
    print(f"[gpu] 5000 steps: max displacement {disp:.4g}, FP32 deviation {dev:.3g} ({dev / disp:.3g} of it), "
          f"velocity {dU:.3g} of the largest speed", flush=True)
    assert disp > 0.1
    assert dev <= 2e-5 * disp
    assert dU <= 1e-4
