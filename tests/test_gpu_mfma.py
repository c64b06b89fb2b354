"""The FP16-stored MRT step with the forward moment transform on the matrix
cores (transform = "mfma", mfma.hpp) against the same step on the vector ALUs
(transform = "valu") and against the FP64 path.  The two FP16 forms differ
only in the order of the transforms' FP32 sums (and the stochastic rounding
that follows), so their fields agree to FP32 rounding amplified by the
occasional flipped rounding of a store; both hold the FP16 bound against FP64.
Rows of 70 cells leave a partial wavefront per row (the VALU path there)."""
import numpy as np
import pytest

from mglc_amd.lbm import Cavity, CavityConfig

pytestmark = pytest.mark.gpu


def run(shape, storage, transform, steps, **kw):
    nx, ny, nz = shape
    cfg = CavityConfig(nx, ny, nz, Reynolds=400.0, U0=0.1, collision="mrt", storage=storage, transform=transform,
                       **kw)
    with Cavity(cfg) as c:
        c.initial()
        c.step(steps)
        return np.stack(c.macro()[1:]), float(c.macro()[0].sum())


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("shape", [(64, 40, 36), (70, 33, 30)], ids=["64x40x36", "70x33x30"])
@pytest.mark.parametrize("loop", [0, 1], ids=["walls", "loopback"])
def test_mfma_transforms_match_valu_and_fp64(shape, loop):
    steps = 600
    print(f"[gpu] mfma {shape} loop={loop}", flush=True)
    kw = dict(halo_loopback=loop)
    v, mv = run(shape, "f16", "valu", steps, **kw)
    m, mm = run(shape, "f16", "mfma", steps, **kw)
    r, mr = run(shape, "f64", "auto", steps, **kw)
    d_mv, d_mr, d_vr = rel_l2(m, v), rel_l2(m, r), rel_l2(v, r)
    print(f"mfma vs valu {d_mv:.3e}, mfma vs f64 {d_mr:.3e}, valu vs f64 {d_vr:.3e}, "
          f"mass {abs(mm - mr) / mr:.2e}", flush=True)
    assert d_mv <= 1e-3
    assert d_mr <= 2e-3 and d_vr <= 2e-3
    assert abs(mm - mr) / mr <= 1e-6


def test_mfma_with_a_halo_face_matches_valu():
    """x faces looped back (periodic x): the x-face lanes of the matrix-core
    waves take their pulls from the halo."""
    shape, steps = (64, 24, 20), 300
    kw = dict(halo_faces=3, halo_loopback=1)
    v, _ = run(shape, "f16", "valu", steps, **kw)
    m, _ = run(shape, "f16", "mfma", steps, **kw)
    assert rel_l2(m, v) <= 1e-3
