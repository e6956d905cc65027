"""Host-side logic of bench.py (no GPU): the one-GPU guard, the traffic
figures it reads from the committed counter summary, and the CPU-baseline
thread count."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_traffic_figures_come_from_the_counter_summary():
    b = _bench()
    # bytes per k_update launch, as profiles/update_kernel_traffic.md states them
    for key, prec in (("particle:1005669x2048", "fp64"), ("particle:1005669x2048", "fp32"),
                      ("cylinder:101753x512", "fp64"), ("cylinder:101753x512", "fp32")):
        t = b._traffic_from_profiles(key, prec)
        assert t is not None and t > 0, (key, prec)
    # above the algorithmic bytes (one read and one write of f), below twice them
    alg = 2 * 8 * 3 * 1005669 * 2048
    t = b._traffic_from_profiles("particle:1005669x2048", "fp64")
    assert alg < t < 2 * alg
    assert b._traffic_from_profiles("particle:1x1", "fp64") is None


def test_cpu_baseline_threads(monkeypatch):
    b = _bench()
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    assert b.host_threads() == 3
    monkeypatch.setenv("OMP_NUM_THREADS", "not a number")
    assert 1 <= b.host_threads() <= 16
