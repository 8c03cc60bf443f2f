"""CPU checks of bench.py's host logic: the dominant kernel class and its
roofline figure, HBM traffic figures used only for the workload they were
measured on, and the one-GPU contract (no device is touched)."""
import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_roofline_takes_the_class_with_most_device_time(bench):
    stats = {"spmm_dot_f32": dict(count=10, ms=20.0, bytes=80e9, flops=1e12),
             "axpby": dict(count=100, ms=5.0, bytes=30e9, flops=0.0),
             "copy": dict(count=0, ms=0.0, bytes=0.0, flops=0.0)}
    name, r = bench.roofline(stats, "mixed", "chung_lu 2000000")
    assert name == "spmm_dot_f32"
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    # 80 GB over 20 ms = 4000 GB/s; 10 launches of 8 GB, 2 ms each
    assert r["achieved"] == pytest.approx(4000.0) and r["frac"] == pytest.approx(0.5)
    assert r["launches"] == 10 and r["avg_ms"] == pytest.approx(2.0) and r["bytes_per_launch"] == 8e9
    assert r["traffic"] is None  # no counter run for this workload


def test_roofline_of_an_mfma_class_is_in_tflops(bench):
    stats = {"gram_mfma_f64": dict(count=4, ms=40.0, bytes=1e9, flops=2e12)}
    name, r = bench.roofline(stats, "mixed", "vardiff3d 160")
    assert name == "gram_mfma_f64"
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s" and r["peak"] == 78.6
    assert r["achieved"] == pytest.approx(50.0) and r["frac"] == pytest.approx(50.0 / 78.6, rel=1e-3)


def test_traffic_only_for_the_profiled_workload(bench):
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        wl = json.load(f)["workload"]
    assert bench.load_traffic("mixed", wl), "traffic.json has the profiled workload's figures"
    assert bench.load_traffic("mixed", "chung_lu 2000000") == {}
    assert bench.load_traffic("fp64", "lap2d 128") == {}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "error" in line and line["n_gpus"] == 2
