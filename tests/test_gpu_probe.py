"""lbm_bandwidth_probe on the GPU: both traffic patterns return a plausible
rate for every storage (at 96^3 the two lattices fit in the last-level
cache, so rates above the HBM peak are possible)."""
import pytest

from mglc_amd.lbm import bandwidth_probe

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_probe_rates_are_plausible(storage):
    print(f"[gpu] bandwidth probe {storage} 96^3", flush=True)
    copy = bandwidth_probe(0, storage, 96, reps=5)
    pull = bandwidth_probe(1, storage, 96, reps=5)
    assert 50.0 < copy < 20000.0
    assert 50.0 < pull < 20000.0


def test_probe_buoyancy_pattern():
    print("[gpu] bandwidth probe buoyancy pattern 96^3", flush=True)
    g = bandwidth_probe(2, "f64", 96, reps=5)
    assert 50.0 < g < 20000.0
