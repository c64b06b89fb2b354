"""CPU tests of bench.py's command-line contract that need no GPU: one GPU
only, and the roofline fields computed from bytes and time."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_roofline_fields():
    sys.path.insert(0, ROOT)
    import bench
    r = bench.roofline(5.1e9, 1.0, traffic=5.2e9)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["achieved"] - 5100.0) < 1e-6
    assert abs(r["frac"] - 5100.0 / 8000.0) < 1e-4
    assert r["traffic"] == 5.2e9


def test_host_cpu_reports_model_and_thread_count(monkeypatch):
    """The CPU legs state what they ran on: the CPU model, the machine's CPU
    count, the CPUs this process may use, and a default thread count capped
    by OMP_NUM_THREADS (a shared box's CPU share)."""
    sys.path.insert(0, ROOT)
    import bench
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    h = bench.host_cpu()
    assert isinstance(h["model"], str) and h["model"]
    assert h["cpu_count"] >= h["usable_cpus"] >= 1
    assert h["threads"] == min(3, h["usable_cpus"])
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_cpu()["threads"] == h["usable_cpus"]
