"""GPU tests of checkpoint / restart (lbm_save_state, lbm_load_state): a run
restarted from a saved state continues bit for bit, for each storage, with
the halo loopback and on the buoyancy path (Coriolis on), and a state is
refused by a context of another shape or kind."""
import numpy as np
import pytest

from mglc_amd.lbm import Cavity, CavityConfig, LbmError, ThermalConfig

pytestmark = pytest.mark.gpu

LBM_ERR_ARG, LBM_ERR_UNSUPPORTED = 1, 7


def fields(c, thermal=False):
    out = list(c.macro())
    if thermal:
        out.append(c.temperature())
    return out


def restart_matches(cfg, first, second, start="initial", thermal=False):
    print(f"[gpu] restart {type(cfg).__name__} {cfg.nx}x{cfg.ny}x{cfg.nz} {first}+{second}", flush=True)
    with Cavity(cfg) as c:
        getattr(c, start)()
        c.step(first)
        blob = c.save_state()
        c.step(second)
        ref = fields(c, thermal)
        ref_check = c.check_thermal() if thermal else c.check()
        steps_ref = c.info().steps_done
    with Cavity(cfg) as d:
        d.load_state(blob)
        assert d.info().steps_done == first
        d.step(second)
        got = fields(d, thermal)
        got_check = d.check_thermal() if thermal else d.check()
        assert d.info().steps_done == steps_ref
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    assert got_check == ref_check


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_cavity_restart_is_exact(storage):
    restart_matches(CavityConfig(40, 36, 32, Reynolds=400, storage=storage), 90, 70)


def test_cavity_restart_with_halo_loopback_is_exact():
    # x and y faces exchange (periodic through the loopback); the z walls and
    # the lid keep a flow going, so the residual is defined
    restart_matches(CavityConfig(32, 30, 28, Reynolds=300, storage="f32", halo_faces=0x0F, halo_loopback=1), 50, 40)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_thermal_restart_is_exact(storage):
    cfg = ThermalConfig(20, 24, 22, Rayleigh=1e5, Ekman=0.001, storage=storage)
    restart_matches(cfg, 300, 200, thermal=True)


def test_state_is_refused_by_another_context():
    with Cavity(CavityConfig(24, 24, 24)) as c:
        c.initial()
        c.step(3)
        blob = c.save_state()
    for other in (CavityConfig(24, 24, 20), CavityConfig(24, 24, 24, storage="f32"),
                  CavityConfig(24, 24, 24, collision="bgk")):
        with Cavity(other) as d:
            with pytest.raises(LbmError) as e:
                d.load_state(blob)
            assert e.value.status == LBM_ERR_ARG
    with Cavity(CavityConfig(24, 24, 24)) as d:
        with pytest.raises(LbmError) as e:
            d.load_state(blob[:-8])
        assert e.value.status == LBM_ERR_ARG


def test_state_keeps_the_fp16_scale_and_halo_flags():
    """An FP16 state holds (f - w) 2^k: a context with another k would misread
    every population by 2^(k1 - k2), so it refuses the state; so does a context
    whose halo loopback or split step differs from the saved one's."""
    with Cavity(CavityConfig(24, 24, 24, storage="f16", f16_scale_log2=12)) as c:
        c.initial()
        c.step(3)
        blob = c.save_state()
    with Cavity(CavityConfig(24, 24, 24, storage="f16", f16_scale_log2=10)) as d:
        with pytest.raises(LbmError) as e:
            d.load_state(blob)
        assert e.value.status == LBM_ERR_ARG and "2^12" in str(e.value)
    # k = 0 selects the default 12: the same scale, accepted
    with Cavity(CavityConfig(24, 24, 24, storage="f16", f16_scale_log2=0)) as d:
        d.load_state(blob)
    with Cavity(CavityConfig(24, 24, 24, storage="f16", f16_scale_log2=12, halo_loopback=1)) as d:
        with pytest.raises(LbmError) as e:
            d.load_state(blob)
        assert e.value.status == LBM_ERR_ARG and "halo_loopback" in str(e.value)


def test_contexts_with_spheres_are_not_checkpointed():
    with Cavity(CavityConfig(24, 24, 24, U0=0.0, tauf=0.8, meq12_rho=1)) as c:
        c.particles_init(np.array([[12.0, 12.0, 12.0]]), np.array([3.0]))
        with pytest.raises(LbmError) as e:
            c.save_state()
        assert e.value.status == LBM_ERR_UNSUPPORTED
