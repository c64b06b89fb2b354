"""GPU tests of the subdomain exchange: six-face + twelve-edge pack/unpack
(mpi_3d_blocked/ex_sendrecv.f90:12-123 layout), the looped-back exchange, and
the decomposition invariance the reference shows across MPI ranks."""
import numpy as np
import pytest

from mglc_amd.hip import DeviceBuffer
from mglc_amd.lbm import FACE_XM, FACE_XP, FACE_ZM, FACE_ZP, Cavity, CavityConfig, LbmError

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def shear_field(nx, ny, nz, amp=0.02, seed=3):
    """A smooth periodic velocity field (Taylor-Green in x-y plus a z shear)
    with small seeded noise, uniform density."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky, kz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    rng = np.random.default_rng(seed)
    u = amp * np.sin(kx * (x + 0.5)) * np.cos(ky * (y + 0.5)) + 1e-4 * rng.standard_normal(x.shape)
    v = -amp * np.cos(kx * (x + 0.5)) * np.sin(ky * (y + 0.5)) + 1e-4 * rng.standard_normal(x.shape)
    w = 0.5 * amp * np.sin(kz * (z + 0.5)) + 1e-4 * rng.standard_normal(x.shape)
    return np.ones_like(u), u, v, w


def test_layout_regions():
    nx, ny, nz = 20, 12, 8
    with Cavity(CavityConfig(nx, ny, nz, tauf=0.6)) as c:
        counts = [c.halo_layout(r)[1] for r in range(18)]
        offs = [c.halo_layout(r)[0] for r in range(18)]
        assert counts[:6] == [5 * ny * nz] * 2 + [5 * nx * nz] * 2 + [5 * nx * ny] * 2
        # edges 7..10 run along z, 11..14 along y, 15..18 along x
        assert counts[6:] == [nz] * 4 + [ny] * 4 + [nx] * 4
        assert offs == list(np.cumsum([0] + counts[:-1]))
        assert c.halo_buffer_bytes() == 8 * sum(counts)
        with pytest.raises(LbmError):
            c.halo_layout(18)


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_loopback_with_walls_changes_nothing(storage):
    """Config of the precision study: the exchange runs every step, looped back,
    while all six faces stay walls -- the walls win, as in the reference where
    bounceback overwrites what streaming pulled from the halo."""
    base = dict(nx=40, ny=36, nz=32, tauf=0.6, storage=storage)
    out = []
    for lb in (0, 1):
        with Cavity(CavityConfig(halo_loopback=lb, **base)) as c:
            c.initial()
            c.step(40)
            out.append(c.macro())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("coll", ["mrt", "bgk"])
@pytest.mark.parametrize("mask,faces", [(7, 63), (1, 3), (6, 60), (4, 48)],
                         ids=["xyz-periodic", "x-periodic", "yz-periodic", "z-periodic"])
def test_periodic_loopback_matches_oracle(oracle_lib, coll, mask, faces):
    """Faces in halo_faces read the looped-back halo: periodic along those axes.
    The oracle's periodic streaming is what a periodic Cartesian exchange of
    the reference would deliver; walls (and the lid, when z+ is a wall) stay."""
    O = oracle_lib
    nx, ny, nz, tau, steps = 24, 20, 18, 0.58, 60
    fields = shear_field(nx, ny, nz)
    p = O.cavity_params(nx, ny, nz, tau, 0.1, collision=O.MRT if coll == "mrt" else O.BGK, periodic_mask=mask)
    ref = O.cavity_run_from(p, steps, *fields)
    print(f"[gpu] periodic mask={mask} {coll}", flush=True)
    with Cavity(CavityConfig(nx, ny, nz, tauf=tau, U0=0.1, collision=coll, halo_faces=faces,
                             halo_loopback=1)) as c:
        c.load_fields(*fields)
        c.step(steps)
        got = c.macro()
    for g, r, name in zip(got, ref, ("rho", "u", "v", "w")):
        assert relmax(g, r) <= TOL, name


@pytest.mark.parametrize("shape", [(65, 3, 4), (2, 2, 2), (33, 17, 5), (3, 64, 2)],
                         ids=["65x3x4", "2x2x2", "33x17x5", "3x64x2"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_periodic_loopback_edge_shapes(oracle_lib, shape, storage):
    """The step kernel's own looped-back stores at edge shapes: a row one lane
    past a wavefront, two-node boxes (both faces of an axis on the same cell
    layer pair), rows shorter than a wavefront holding both x-face lanes.
    FP64 against the oracle's periodic streaming; FP32 against FP64."""
    O = oracle_lib
    nx, ny, nz = shape
    tau, steps = 0.62, 40
    fields = shear_field(nx, ny, nz)
    kw = dict(tauf=tau, U0=0.1, halo_faces=63, halo_loopback=1)
    print(f"[gpu] periodic edge shape {shape} {storage}", flush=True)
    with Cavity(CavityConfig(nx, ny, nz, storage=storage, **kw)) as c:
        c.load_fields(*fields)
        c.step(steps)
        got = c.macro()
    if storage == "f64":
        # relative to the flow's scales: rho ~ 1, velocities ~ the initial
        # amplitude 0.02 (on these small boxes the fields decay far below it,
        # where a bound relative to their own maximum would measure rounding)
        ref = O.cavity_run_from(O.cavity_params(nx, ny, nz, tau, 0.1, periodic_mask=7), steps, *fields)
        assert relmax(got[0], ref[0]) <= TOL, "rho"
        for g, r, name in zip(got[1:], ref[1:], ("u", "v", "w")):
            assert np.max(np.abs(g - r)) <= TOL * 0.02, name
    else:
        with Cavity(CavityConfig(nx, ny, nz, storage="f64", **kw)) as c:
            c.load_fields(*fields)
            c.step(steps)
            ref = c.macro()
        for g, r, name in zip(got[1:], ref[1:], ("u", "v", "w")):
            assert np.max(np.abs(g - r)) <= 1e-5, name


def test_fully_periodic_conserves_mass_and_momentum():
    nx, ny, nz = 32, 24, 16
    fields = shear_field(nx, ny, nz, amp=0.05)
    with Cavity(CavityConfig(nx, ny, nz, tauf=0.7, U0=0.0, halo_faces=63, halo_loopback=1)) as c:
        c.load_fields(*fields)
        c.step(200)
        rho, u, v, w = c.macro()
    rho0, u0, v0, w0 = fields
    assert abs(rho.sum() - rho0.sum()) / rho0.sum() < 1e-13
    for a, b in ((u, u0), (v, v0), (w, w0)):
        assert abs((rho * a).sum() - (rho0 * b).sum()) < 1e-11 * rho0.size


def test_taylor_green_decay_rate():
    """Textbook pin of the periodic path: a 2-D Taylor-Green vortex in a
    periodic box decays as exp(-nu (kx^2 + ky^2) t), nu = (tau - 1/2)/3 (BGK)."""
    n, tau, steps, amp = 64, 0.8, 400, 0.01
    z, y, x = np.meshgrid(np.arange(4), np.arange(n), np.arange(n), indexing="ij")
    k = 2 * np.pi / n
    u = amp * np.sin(k * x) * np.cos(k * y)
    v = -amp * np.cos(k * x) * np.sin(k * y)
    rho = 1.0 - 3 * amp ** 2 / 4 * (np.cos(2 * k * x) + np.cos(2 * k * y))
    with Cavity(CavityConfig(n, n, 4, tauf=tau, U0=0.0, collision="bgk", halo_faces=63, halo_loopback=1)) as c:
        c.load_fields(rho, u, v, np.zeros_like(u))
        c.step(steps)
        _, u1, v1, _ = c.macro()
    nu = (tau - 0.5) / 3.0
    expect = np.exp(-nu * 2 * k * k * steps)
    got = float(np.sum(u1 * u) / np.sum(u * u))
    assert got == pytest.approx(expect, rel=5e-3)


def exchange(ctxs, sends, recvs, links):
    """Pack every subdomain, then copy each linked region of the sender's packed
    buffer into the receiver's incoming region (the transport an MPI or RCCL
    exchange would do), then unpack."""
    for c, s in zip(ctxs, sends):
        c.halo_pack(s.address)
    for (dst, dst_region, src, src_region) in links:
        doff, dcnt = ctxs[dst].halo_layout(dst_region)
        soff, scnt = ctxs[src].halo_layout(src_region)
        assert dcnt == scnt
        elem = ctxs[dst].halo_buffer_bytes() // sum(ctxs[dst].halo_layout(r)[1] for r in range(18))
        recvs[dst].copy_from(sends[src], doff * elem, soff * elem, dcnt * elem)
    for c, r in zip(ctxs, recvs):
        c.halo_unpack(r.address)


@pytest.mark.parametrize("axis", ["x", "z"])
def test_two_subdomains_equal_one_domain(axis):
    """Decomposition invariance: a cavity split in two subdomains that exchange
    halos every step reproduces the undivided cavity exactly (the reference's
    1/2/4/8-rank runs agree the same way).  Splitting along z puts the lid on
    the upper subdomain only."""
    nx, ny, nz, tau, steps = 32, 20, 24, 0.6, 50
    with Cavity(CavityConfig(nx, ny, nz, tauf=tau)) as whole:
        whole.initial()
        whole.step(steps)
        ref = whole.macro()
    if axis == "x":
        h = nx // 2
        cfgs = [CavityConfig(h, ny, nz, tauf=tau, halo_faces=1 << FACE_XP),
                CavityConfig(nx - h, ny, nz, tauf=tau, halo_faces=1 << FACE_XM)]
        links = [(0, FACE_XP, 1, FACE_XM), (1, FACE_XM, 0, FACE_XP)]
    else:
        h = nz // 2
        cfgs = [CavityConfig(nx, ny, h, tauf=tau, halo_faces=1 << FACE_ZP),
                CavityConfig(nx, ny, nz - h, tauf=tau, halo_faces=1 << FACE_ZM)]
        links = [(0, FACE_ZP, 1, FACE_ZM), (1, FACE_ZM, 0, FACE_ZP)]
    ctxs = [Cavity(c) for c in cfgs]
    sends = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    recvs = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    try:
        for c in ctxs:
            c.initial()  # u = U0 only where z+ is the lid wall, as initial.f90:54-61
        for _ in range(steps):
            exchange(ctxs, sends, recvs, links)
            for c in ctxs:
                c.step(1)
        parts = [c.macro() for c in ctxs]
        cat = 2 if axis == "x" else 0
        got = [np.concatenate([parts[0][q], parts[1][q]], axis=cat) for q in range(4)]
        for g, r, name in zip(got, ref, ("rho", "u", "v", "w")):
            assert np.array_equal(g, r), (name, relmax(g, r))
    finally:
        for b in sends + recvs:
            b.close()
        for c in ctxs:
            c.close()


def test_external_halo_requires_single_steps():
    with Cavity(CavityConfig(16, 16, 16, tauf=0.6, halo_faces=1)) as c:
        c.initial()
        with pytest.raises(LbmError) as ei:
            c.step(2)
        assert ei.value.status == 6
