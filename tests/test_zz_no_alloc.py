"""lbm_check_sums_noalloc (the residual sums over caller device arrays, with
its partial sums in the library's fixed per-device array) against
lbm_check_sums (the same kernels on a context's own buffer), bit for bit, on
two calls in a row, the second with a smaller input; both through the C-ABI."""
import ctypes

import numpy as np
import pytest

from mglc_amd import hip
from mglc_amd.lbm import Cavity, CavityConfig

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def upload(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    buf = hip.DeviceBuffer(a.nbytes)
    rc = hip.runtime().hipMemcpy(buf.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D)
    assert rc == 0
    return buf


def download(buf, n):
    out = np.empty(n, dtype=np.float64)
    rc = hip.runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), buf.ptr, out.nbytes, D2H)
    assert rc == 0
    return out


def sums_both_ways(lib, shape, steps):
    """A cavity after `steps` steps: (e1, e2) from lbm_check_sums on the
    context, and from lbm_check_sums_noalloc on device copies of the same
    u, v, w with up = vp = 0 (a context's previous fields after lbm_initial),
    plus the up / vp the latter left."""
    nx, ny, nz = shape
    print(f"[gpu] no-alloc sums {nx}x{ny}x{nz} after {steps} steps", flush=True)
    with Cavity(CavityConfig(nx, ny, nz, tauf=0.62, collision="mrt")) as c:
        c.initial()
        c.step(steps)
        _, u, v, w = c.macro()
        ref = c.check_sums()
    n = u.size
    bufs = [upload(a.ravel()) for a in (u, v, w, np.zeros(n), np.zeros(n))]
    sums = upload(np.full(2, np.nan))
    try:
        rc = lib.lbm_check_sums_noalloc(*(b.ptr for b in bufs), n, sums.ptr, None)
        assert rc == 0, lib.lbm_last_error()
        assert hip.runtime().hipDeviceSynchronize() == 0
        got = download(sums, 2)
        up, vp = download(bufs[3], n), download(bufs[4], n)
    finally:
        for b in bufs + [sums]:
            b.close()
    return ref, tuple(got), (u.ravel(), v.ravel()), (up, vp)


def test_noalloc_sums_equal_the_context_sums(hip_lib):
    for shape, steps in (((70, 44, 36), 40), ((17, 9, 12), 25)):
        ref, got, (u, v), (up, vp) = sums_both_ways(hip_lib, shape, steps)
        assert ref[0] > 0 and ref[1] > 0
        assert got == ref, (shape, got, ref)
        assert np.array_equal(up, u) and np.array_equal(vp, v)
