"""The split step (halo_overlap): boundary shell and interior as the
reference's non-blocking code computes them (lid3_mpi_nonblock.f90:1108-1231),
with the looped-back exchange on a second stream, and the begin / end API a
caller uses to overlap its own transport.  Every cell is updated by the same
code as in the fused step, so results must be identical bit for bit."""
import numpy as np
import pytest

from mglc_amd.hip import DeviceBuffer
from mglc_amd.lbm import FACE_YM, FACE_YP, FACE_ZM, FACE_ZP, Cavity, CavityConfig, LbmError

pytestmark = pytest.mark.gpu


def shear(nx, ny, nz, seed=5):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(seed)
    u = 0.03 * np.sin(2 * np.pi * (y + 0.5) / ny) * np.cos(2 * np.pi * (z + 0.5) / nz)
    v = 0.02 * np.cos(2 * np.pi * (x + 0.5) / nx) + 1e-4 * rng.standard_normal(x.shape)
    w = 0.02 * np.sin(2 * np.pi * (y + 0.5) / ny) + 1e-4 * rng.standard_normal(x.shape)
    return np.ones_like(u), u, v, w


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
@pytest.mark.parametrize("collision", ["mrt", "bgk"])
@pytest.mark.parametrize("mask", [0x3C, 0x0C, 0x30])
def test_split_loopback_equals_fused(storage, collision, mask):
    nx, ny, nz = 70, 22, 18
    print(f"[gpu] split loopback {storage} {collision} mask 0x{mask:x}", flush=True)
    fields = shear(nx, ny, nz)
    kw = dict(tauf=0.62, U0=0.05, storage=storage, collision=collision, halo_faces=mask, halo_loopback=1)
    with Cavity(CavityConfig(nx, ny, nz, **kw)) as fused, \
            Cavity(CavityConfig(nx, ny, nz, halo_overlap=1, **kw)) as split:
        fused.load_fields(*fields)
        split.load_fields(*fields)
        for chunk in (1, 5, 31):
            fused.step(chunk)
            split.step(chunk)
            a, b = fused.macro(), split.macro()
            for q, name in enumerate(("rho", "u", "v", "w")):
                assert np.array_equal(a[q], b[q]), (chunk, name)
        assert fused.info().steps_done == split.info().steps_done == 37


def exchange_links(ctxs, sends, recvs, links):
    for (dst, dst_region, src, src_region) in links:
        doff, dcnt = ctxs[dst].halo_layout(dst_region)
        soff, scnt = ctxs[src].halo_layout(src_region)
        assert dcnt == scnt
        elem = ctxs[dst].halo_buffer_bytes() // sum(ctxs[dst].halo_layout(r)[1] for r in range(18))
        recvs[dst].copy_from(sends[src], doff * elem, soff * elem, dcnt * elem)


@pytest.mark.parametrize("axis", ["y", "z"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_split_api_two_subdomains_equal_one_domain(axis, storage):
    """begin (shell) -> pack -> transport -> end (interior) -> unpack, per step,
    on two subdomains, reproduces the undivided cavity exactly.  (FP16 storage
    rounds stochastically with a stream keyed on the cell index within the
    subdomain, so a decomposed FP16 run is a different, equally valid,
    rounding sequence; its split is checked against the fused step above.)"""
    nx, ny, nz, tau, steps = 40, 24, 20, 0.6, 30
    print(f"[gpu] split api {axis} {storage}", flush=True)
    with Cavity(CavityConfig(nx, ny, nz, tauf=tau, storage=storage)) as whole:
        whole.initial()
        whole.step(steps)
        ref = whole.macro()
    if axis == "y":
        h = ny // 2
        cfgs = [CavityConfig(nx, h, nz, tauf=tau, storage=storage, halo_faces=1 << FACE_YP, halo_overlap=1),
                CavityConfig(nx, ny - h, nz, tauf=tau, storage=storage, halo_faces=1 << FACE_YM, halo_overlap=1)]
        links = [(0, FACE_YP, 1, FACE_YM), (1, FACE_YM, 0, FACE_YP)]
    else:
        h = nz // 2
        cfgs = [CavityConfig(nx, ny, h, tauf=tau, storage=storage, halo_faces=1 << FACE_ZP, halo_overlap=1),
                CavityConfig(nx, ny, nz - h, tauf=tau, storage=storage, halo_faces=1 << FACE_ZM, halo_overlap=1)]
        links = [(0, FACE_ZP, 1, FACE_ZM), (1, FACE_ZM, 0, FACE_ZP)]
    ctxs = [Cavity(c) for c in cfgs]
    sends = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    recvs = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    try:
        for c in ctxs:
            c.initial()
        # the halo of the initial state
        for c, s in zip(ctxs, sends):
            c.halo_pack(s.address)
        exchange_links(ctxs, sends, recvs, links)
        for c, r in zip(ctxs, recvs):
            c.halo_unpack(r.address)
        for _ in range(steps):
            for c in ctxs:
                c.step_begin()
            for c, s in zip(ctxs, sends):
                c.halo_pack(s.address)        # the shell of the step in progress
            exchange_links(ctxs, sends, recvs, links)
            for c in ctxs:
                c.step_end()
            for c, r in zip(ctxs, recvs):
                c.halo_unpack(r.address)
        parts = [c.macro() for c in ctxs]
        cat = 1 if axis == "y" else 0
        got = [np.concatenate([parts[0][q], parts[1][q]], axis=cat) for q in range(4)]
        for g, r, name in zip(got, ref, ("rho", "u", "v", "w")):
            assert np.array_equal(g, r), name
    finally:
        for b in sends + recvs:
            b.close()
        for c in ctxs:
            c.close()


def test_split_api_state_errors():
    with Cavity(CavityConfig(32, 16, 16, tauf=0.6)) as plain:
        plain.initial()
        with pytest.raises(LbmError) as ei:
            plain.step_begin()
        assert ei.value.status == 6
    with Cavity(CavityConfig(32, 16, 16, tauf=0.6, halo_faces=1 << FACE_ZP, halo_overlap=1)) as c:
        c.initial()
        with pytest.raises(LbmError):
            c.step_end()
        c.step_begin()
        for call in (lambda: c.step(1), lambda: c.check(), lambda: c.macro(), lambda: c.step_begin(),
                     lambda: c.save_state()):
            with pytest.raises(LbmError) as ei:
                call()
            assert ei.value.status == 6
        c.step_end()
        assert c.info().steps_done == 1
        c.check()


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_split_loopback_restart_is_exact(storage):
    """A looped-back split context saved mid-run and restored in a fresh
    context continues bit for bit (the restored halo is re-exchanged from the
    restored populations before the first step)."""
    nx, ny, nz = 64, 20, 16
    print(f"[gpu] split loopback restart {storage}", flush=True)
    kw = dict(tauf=0.6, U0=0.05, storage=storage, halo_faces=0x3C, halo_loopback=1, halo_overlap=1)
    fields = shear(nx, ny, nz, seed=9)
    with Cavity(CavityConfig(nx, ny, nz, **kw)) as a:
        a.load_fields(*fields)
        a.step(12)
        blob = a.save_state()
        a.step(7)
        a.step(6)
        want = a.macro()
    with Cavity(CavityConfig(nx, ny, nz, **kw)) as b:
        b.load_state(blob)
        b.step(13)
        got = b.macro()
    for q in range(4):
        assert np.array_equal(got[q], want[q]), q


def test_split_loopback_after_a_caller_driven_step():
    """A step made with step_begin / step_end on a looped-back split context
    leaves the halo to the caller; the next lbm_step exchanges it again from
    the state, so the run equals an uninterrupted one."""
    nx, ny, nz = 64, 20, 16
    print("[gpu] split loopback after begin/end", flush=True)
    kw = dict(tauf=0.6, U0=0.05, storage="f64", halo_faces=0x3C, halo_loopback=1, halo_overlap=1)
    fields = shear(nx, ny, nz, seed=4)
    with Cavity(CavityConfig(nx, ny, nz, **kw)) as a, Cavity(CavityConfig(nx, ny, nz, **kw)) as b:
        a.load_fields(*fields)
        b.load_fields(*fields)
        a.step(5)
        a.step_begin()
        a.step_end()
        a.step(4)
        b.step(10)
        for q, (x, y) in enumerate(zip(a.macro(), b.macro())):
            assert np.array_equal(x, y), q
