"""GPU tests of the buoyancy path's subdomain exchange: the D3Q19 regions of
the isothermal path plus one D3Q7 region per face (message_exchange.F90:
f_post's 5 populations per face and 1 per edge line, g_post's 1 per face).
Split cavities that exchange every step reproduce the undivided cavity bit
for bit; the looped-back exchange equals the caller-driven one; a periodic
axis keeps an x-uniform flow x-uniform."""
import numpy as np
import pytest

from mglc_amd.hip import DeviceBuffer
from mglc_amd.lbm import (FACE_XM, FACE_XP, FACE_YM, FACE_YP, FACE_ZM, FACE_ZP, Cavity, LbmError, ThermalConfig,
                          buoyancy_initial_temperature)

pytestmark = pytest.mark.gpu

G0 = 18  # first D3Q7 face region
ELEM = {"f64": 8, "f32": 4, "f16": 2}


def cfg(nx, ny, nz, storage="f64", **kw):
    # side-heated cavity with rotation (Coriolis on), N fixed so every
    # subdomain has the whole cavity's parameters
    return ThermalConfig(nx, ny, nz, Rayleigh=1e5, Prandtl=0.71, Thot=0.5, Tcold=-0.5, thermal_axis=1,
                         Ekman=0.001, N=24, storage=storage, **kw)


def start_fields(nx, ny, nz, seed=2024):
    T = buoyancy_initial_temperature(nx, ny, nz, seed=seed)
    z = np.zeros_like(T)
    return np.ones_like(T), z, z, z, T


def exchange(ctxs, sends, recvs, links, storage):
    """Pack every subdomain; copy each linked region of the sender's packed
    buffer into the receiver's incoming region (an MPI / RCCL transport);
    unpack."""
    e = ELEM[storage]
    for c, s in zip(ctxs, sends):
        c.halo_pack(s.address)
    for (dst, dreg, src, sreg) in links:
        doff, dcnt = ctxs[dst].halo_layout(dreg)
        soff, scnt = ctxs[src].halo_layout(sreg)
        assert dcnt == scnt
        recvs[dst].copy_from(sends[src], doff * e, soff * e, dcnt * e)
    for c, r in zip(ctxs, recvs):
        c.halo_unpack(r.address)


def face_links(a, fa, b, fb):
    """Subdomain a's face fa borders subdomain b's face fb: the D3Q19 face
    regions and the D3Q7 face regions, both ways."""
    return [(a, fa, b, fb), (b, fb, a, fa), (a, G0 + fa, b, G0 + fb), (b, G0 + fb, a, G0 + fa)]


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_two_thermal_subdomains_equal_one_domain(axis, storage):
    nx, ny, nz, steps = 24, 22, 20, 120
    fields = start_fields(nx, ny, nz)
    with Cavity(cfg(nx, ny, nz, storage)) as whole:
        whole.load_fields(*fields)
        whole.step(steps)
        ref = (*whole.macro(), whole.temperature())
    ax = {"x": 2, "y": 1, "z": 0}[axis]
    n = (nz, ny, nx)[ax]
    h = n // 2
    lo_face, hi_face = {"x": (FACE_XM, FACE_XP), "y": (FACE_YM, FACE_YP), "z": (FACE_ZM, FACE_ZP)}[axis]
    sl = [[slice(None)] * 3 for _ in range(2)]
    sl[0][ax], sl[1][ax] = slice(0, h), slice(h, n)
    shapes = [tuple(fields[0][tuple(s)].shape) for s in sl]
    cfgs = [cfg(sh[2], sh[1], sh[0], storage, halo_faces=1 << hi_face) for sh in shapes[:1]] + \
           [cfg(sh[2], sh[1], sh[0], storage, halo_faces=1 << lo_face) for sh in shapes[1:]]
    links = face_links(0, hi_face, 1, lo_face)
    print(f"[gpu] thermal split {axis} {storage}", flush=True)
    ctxs = [Cavity(c) for c in cfgs]
    sends = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    recvs = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    try:
        for c, s in zip(ctxs, sl):
            c.load_fields(*(np.ascontiguousarray(f[tuple(s)]) for f in fields))
        for _ in range(steps):
            exchange(ctxs, sends, recvs, links, storage)
            for c in ctxs:
                c.step(1)
        parts = [(*c.macro(), c.temperature()) for c in ctxs]
        for q, name in enumerate(("rho", "u", "v", "w", "T")):
            got = np.concatenate([parts[0][q], parts[1][q]], axis=ax)
            assert np.array_equal(got, ref[q]), name
    finally:
        for b in sends + recvs:
            b.close()
        for c in ctxs:
            c.close()


def test_four_thermal_subdomains_with_edges_equal_one_domain():
    """A 2 x 2 split across y and z: the diagonal D3Q19 populations crossing
    both cut faces travel through the edge regions."""
    nx, ny, nz, steps = 20, 24, 22, 100
    fields = start_fields(nx, ny, nz, seed=7)
    with Cavity(cfg(nx, ny, nz)) as whole:
        whole.load_fields(*fields)
        whole.step(steps)
        ref = (*whole.macro(), whole.temperature())
    hy, hz = ny // 2, nz // 2
    # piece index = 2 * (upper z) + (upper y)
    pieces = []
    for pz in range(2):
        for py in range(2):
            faces = (1 << (FACE_YP if py == 0 else FACE_YM)) | (1 << (FACE_ZP if pz == 0 else FACE_ZM))
            s = (slice(0, hz) if pz == 0 else slice(hz, nz), slice(0, hy) if py == 0 else slice(hy, ny), slice(None))
            pieces.append((s, faces))
    links = face_links(0, FACE_YP, 1, FACE_YM) + face_links(2, FACE_YP, 3, FACE_YM) + \
        face_links(0, FACE_ZP, 2, FACE_ZM) + face_links(1, FACE_ZP, 3, FACE_ZM)
    # edges: region 6 + (a - 7) for the diagonal population a; a population
    # leaving piece p through its y and z cut faces arrives in the diagonal
    # piece through the same edge region
    ey = {a: e for a, e in zip(range(7, 19), [1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])}
    ez = {a: e for a, e in zip(range(7, 19), [0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])}
    for a in range(15, 19):
        src = (0 if ey[a] > 0 else 1) + 2 * (0 if ez[a] > 0 else 1)  # the piece it leaves through both cuts
        dst = 3 - src                                                  # the diagonal piece
        links.append((dst, 6 + (a - 7), src, 6 + (a - 7)))
    print("[gpu] thermal split 2x2 y-z", flush=True)
    ctxs = [Cavity(cfg(fields[0][s].shape[2], fields[0][s].shape[1], fields[0][s].shape[0], halo_faces=f))
            for s, f in pieces]
    sends = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    recvs = [DeviceBuffer(c.halo_buffer_bytes()) for c in ctxs]
    try:
        for c, (s, _) in zip(ctxs, pieces):
            c.load_fields(*(np.ascontiguousarray(f[s]) for f in fields))
        for _ in range(steps):
            exchange(ctxs, sends, recvs, links, "f64")
            for c in ctxs:
                c.step(1)
        parts = [(*c.macro(), c.temperature()) for c in ctxs]
        for q, name in enumerate(("rho", "u", "v", "w", "T")):
            lower = np.concatenate([parts[0][q], parts[1][q]], axis=1)
            upper = np.concatenate([parts[2][q], parts[3][q]], axis=1)
            assert np.array_equal(np.concatenate([lower, upper], axis=0), ref[q]), name
    finally:
        for b in sends + recvs:
            b.close()
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("storage", ["f64", "f16"])
def test_thermal_loopback_equals_caller_exchange(storage):
    """halo_loopback = 1 (the exchange run inside lbm_step) against the same
    context exchanging with itself through pack, a copy of each x face's
    regions into the opposite face's, and unpack."""
    nx, ny, nz, steps = 20, 18, 16, 60
    fields = start_fields(nx, ny, nz, seed=11)
    faces = (1 << FACE_XM) | (1 << FACE_XP)
    out = []
    with Cavity(cfg(nx, ny, nz, storage, halo_faces=faces, halo_loopback=1)) as c:
        c.load_fields(*fields)
        c.step(steps)
        out.append((*c.macro(), c.temperature()))
    links = face_links(0, FACE_XM, 0, FACE_XP)
    with Cavity(cfg(nx, ny, nz, storage, halo_faces=faces)) as c:
        s, r = DeviceBuffer(c.halo_buffer_bytes()), DeviceBuffer(c.halo_buffer_bytes())
        try:
            c.load_fields(*fields)
            for _ in range(steps):
                exchange([c], [s], [r], links, storage)
                c.step(1)
            out.append((*c.macro(), c.temperature()))
        finally:
            s.close()
            r.close()
    for a, b, name in zip(out[0], out[1], ("rho", "u", "v", "w", "T")):
        assert np.array_equal(a, b), name


def test_periodic_axis_keeps_a_uniform_flow_uniform():
    """Periodic along x (both x faces looped back) with a start uniform in x:
    every x column sees the same neighbourhood, so the fields stay uniform in
    x bit for bit while the side-heated flow develops in y-z."""
    nx, ny, nz, steps = 16, 20, 18, 300
    T = np.repeat(buoyancy_initial_temperature(1, ny, nz, seed=5), nx, axis=2)
    z = np.zeros_like(T)
    with Cavity(cfg(nx, ny, nz, halo_faces=(1 << FACE_XM) | (1 << FACE_XP), halo_loopback=1)) as c:
        c.load_fields(np.ones_like(T), z, z, z, T)
        c.step(steps)
        rho, u, v, w = c.macro()
        Tn = c.temperature()
    for f, name in ((rho, "rho"), (u, "u"), (v, "v"), (w, "w"), (Tn, "T")):
        assert np.array_equal(f, np.repeat(f[:, :, :1], nx, axis=2)), name
    assert np.abs(w).max() > 1e-8  # the flow does develop


def test_thermal_halo_layout_and_refusals():
    nx, ny, nz = 12, 10, 8
    with Cavity(cfg(nx, ny, nz, halo_faces=1)) as c:
        counts = [c.halo_layout(r)[1] for r in range(24)]
        assert counts[18:] == [ny * nz] * 2 + [nx * nz] * 2 + [nx * ny] * 2
        assert c.halo_buffer_bytes() == 8 * sum(counts)
        with pytest.raises(LbmError):
            c.halo_layout(24)
    with pytest.raises(LbmError) as ei:
        Cavity(cfg(nx, ny, nz, halo_faces=0x30, halo_overlap=1))
    assert ei.value.status == 7
