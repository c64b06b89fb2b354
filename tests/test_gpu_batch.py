"""GPU tests of the batched cavities: one persistent kernel over many uneven
instances, hipGraph-replayed 100-step blocks, autotuned launch shape.  Every
instance must equal a single-context run of the same parameters exactly, after
every graph replay (with host work and synchronisation in between), for both
the persistent path and the one-launch-per-instance baseline."""
import numpy as np
import pytest

from mglc_amd.lbm import Batch, Cavity, batch_configs

pytestmark = pytest.mark.gpu


def single_runs(cfgs, steps):
    out = []
    for c in cfgs:
        with Cavity(c) as cav:
            cav.initial()
            cav.step(steps)
            out.append(cav.macro())
    return out


@pytest.mark.parametrize("storage", ["f32", "f64", "f16"])
def test_batch_equals_single_contexts_across_graph_replays(storage):
    cfgs = batch_configs(count=6, seed=7, sizes=(16, 24, 32, 40), storage=storage)
    print(f"[gpu] batch {storage} sizes={[c.nx for c in cfgs]}", flush=True)
    with Batch(cfgs) as b:
        b.initial()
        done = 0
        for chunk in (100, 100, 250):        # graph, graph, 2 graphs + 50 direct steps
            b.step(chunk)
            done += chunk
            b.synchronize()
            ref = single_runs(cfgs, done)
            for m in range(len(cfgs)):
                for g, r in zip(b.macro(m), ref[m]):
                    assert np.array_equal(g, r), (storage, m, done)


def test_batch_each_baseline_equals_persistent():
    cfgs = batch_configs(count=5, seed=3, sizes=(16, 32, 48))
    with Batch(cfgs) as b1, Batch(cfgs) as b2:
        b1.initial()
        b2.initial()
        b1.step(130)
        b2.step_each(130)
        for m in range(len(cfgs)):
            for g, r in zip(b1.macro(m), b2.macro(m)):
                assert np.array_equal(g, r)


def test_autotune_keeps_results_and_picks_a_valid_shape():
    cfgs = batch_configs(count=8, seed=11, sizes=(16, 24, 32))
    with Batch(cfgs) as b:
        threads, blocks = b.autotune(trial_steps=2)
        assert threads in (64, 128, 256) and blocks >= 1
        info = b.info()
        assert (info.threads, info.blocks) == (threads, blocks)
        b.step(120)
        ref = single_runs(cfgs, 120)
        for m in range(len(cfgs)):
            for g, r in zip(b.macro(m), ref[m]):
                assert np.array_equal(g, r)


@pytest.mark.parametrize("cpt", [2, 4])
def test_cells_per_thread_change_nothing(cpt):
    """Each wave taking 2 or 4 consecutive tiles per turn: the same cells, the
    same per-cell code, the same bits as one tile per turn."""
    cfgs = batch_configs(count=7, seed=5, sizes=(16, 24, 40))
    with Batch(cfgs) as b1, Batch(cfgs) as b2:
        b2.set_cells_per_thread(cpt)
        assert b2.info().cells_per_thread == cpt
        for b in (b1, b2):
            b.initial()
            b.step(230)
        for m in range(len(cfgs)):
            for g, r in zip(b2.macro(m), b1.macro(m)):
                assert np.array_equal(g, r), (cpt, m)


def test_full_batch_10000_steps_equals_single_contexts():
    """The batched workload itself (64 cavities, N in {32..96}, FP32, autotuned
    launch, 100-step graph replays) for its 10000 steps: one instance of every
    size present equals a single-context run of the same parameters bit for
    bit."""
    cfgs = batch_configs()
    with Batch(cfgs) as b:
        threads, blocks = b.autotune(trial_steps=2)
        info = b.info()
        print(f"[gpu] full batch: {threads} threads x {blocks} blocks, {info.cells_per_thread} cells per thread",
              flush=True)
        b.initial()
        for _ in range(10):
            b.step(1000)
        b.synchronize()
        sample = {}
        for m, c in enumerate(cfgs):
            sample.setdefault(c.nx, m)
        assert sorted(sample) == sorted({c.nx for c in cfgs})
        got = {m: b.macro(m) for m in sample.values()}
    for n, m in sorted(sample.items()):
        print(f"[gpu] single {n}^3 (instance {m}) 10000 steps", flush=True)
        ref = single_runs([cfgs[m]], 10000)[0]
        for g, r in zip(got[m], ref):
            assert np.array_equal(g, r), (n, m)


def test_batch_config_sizes_and_throughput_vs_one_at_a_time():
    """The batched case (64 cavities, N in {32..96}, FP32): the persistent
    graph-replayed path against launching the instances one at a time, on the
    same device memory.  Prints both aggregate MLUPS."""
    from mglc_amd.hip import Event, Stream  # noqa: F401
    import time
    cfgs = batch_configs()
    assert len(cfgs) == 64 and {c.nx for c in cfgs} <= {32, 48, 64, 80, 96}
    with Batch(cfgs) as b:
        b.autotune(trial_steps=3)
        cells = b.info().total_cells
        b.initial()
        b.step(100)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(500)
        b.synchronize()
        t_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        b.step_each(100)
        b.synchronize()
        t_each = (time.perf_counter() - t0) * 5
    mlups_batch = cells * 500 / t_batch / 1e6
    mlups_each = cells * 500 / t_each / 1e6
    print(f"batched persistent+graph: {mlups_batch:.0f} MLUPS; one at a time: {mlups_each:.0f} MLUPS")
    assert mlups_batch > mlups_each
