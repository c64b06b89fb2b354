"""Maximum sizes: lattices of 0.8 and 1.07 G cells (121-161 GB and 133 GB of
HBM with both A-B buffers and the FP64 macro fields), close to the 32-bit cell
index limit, with population offsets far beyond 2^32 elements.  A lid-driven
cavity started at rest changes only within the light cone of the lid (one
cell per step), so after a few steps

* every cell more than `steps` cells below the lid is exactly at rest, and
* a cell near the middle of the lid sees the same inputs as the same cell of a
  small cavity, so (in FP32, deterministic rounding) it holds the same bits.
"""
import ctypes

import numpy as np
import pytest

from mglc_amd.lbm import Cavity, CavityConfig, _dp

pytestmark = pytest.mark.gpu


def download(c, fields):
    """Only the requested macro fields (the others stay on the device)."""
    out = {name: np.empty(c.shape, dtype=np.float64) for name in fields}
    ptrs = [_dp(out[n]) if n in out else None for n in ("rho", "u", "v", "w")]
    c._call("lbm_download_macro", c._ctx, *ptrs)
    return out


def test_fp32_0p8g_cells_light_cone_matches_small_cavity():
    nx, ny, nz, steps = 1024, 1024, 768, 3
    tau = 0.56
    print(f"[gpu] large FP32 {nx}x{ny}x{nz}", flush=True)
    with Cavity(CavityConfig(nx, ny, nz, tauf=tau, storage="f32")) as big:
        assert big.info().device_bytes > 120e9
        big.initial()
        big.step(steps)
        f = download(big, ("rho", "u"))
        rho, u = f["rho"], f["u"]
        # below the light cone: exactly the initial state
        assert np.all(rho[: nz - 1 - steps - 1] == 1.0)
        assert np.all(u[: nz - 1 - steps - 1] == 0.0)
        assert np.isfinite(rho).all() and np.isfinite(u).all()
        # a window around the lid centre, 2 * steps + 1 wide, all within reach of the lid only
        h = steps + 1
        win = (slice(nz - h, nz), slice(ny // 2 - h, ny // 2 + h), slice(nx // 2 - h, nx // 2 + h))
        big_rho, big_u = rho[win].copy(), u[win].copy()
        del rho, u, f
    n = 32
    with Cavity(CavityConfig(n, n, n, tauf=tau, storage="f32")) as small:
        small.initial()
        small.step(steps)
        f = download(small, ("rho", "u"))
        win_s = (slice(n - h, n), slice(n // 2 - h, n // 2 + h), slice(n // 2 - h, n // 2 + h))
        assert np.array_equal(big_rho, f["rho"][win_s])
        assert np.array_equal(big_u, f["u"][win_s])
        assert np.any(big_u != 0.0)


def test_fp16_2p30_cells_rest_region_and_mass():
    n, steps = 1024, 3
    print(f"[gpu] large FP16 {n}^3", flush=True)
    with Cavity(CavityConfig(n, n, n, tauf=0.56, storage="f16")) as c:
        assert c.info().device_bytes > 120e9
        c.initial()
        c.step(steps)
        rho = download(c, ("rho",))["rho"]
        assert np.all(rho[: n - 1 - steps - 1] == 1.0)
        assert np.isfinite(rho).all()
        mass = float(rho.sum(dtype=np.float64))
        assert abs(mass - float(n) ** 3) / float(n) ** 3 < 1e-9
        e = ctypes.c_double()
        c._call("lbm_check", c._ctx, ctypes.byref(e))
        assert np.isfinite(e.value)
