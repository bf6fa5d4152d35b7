"""bench.py's contract pieces that run without a GPU: the one-GPU guard, the
counter summary it reads, the roofline record and the CPU baseline leg."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_counter_summary_parses():
    for prec in ("fp32", "fp32_3m", "fp64"):
        c = bench.counters(prec)
        assert c is not None, prec
        assert c["traffic_bytes"] > c["algorithmic_bytes"] > 0
        assert 0 < c["mfma_util_pct"] < 100


def test_roofline_record():
    flops = 8.0 * (4096 - 256) ** 2 * 128
    info = {"kernel_seconds": 10 * 150e-6, "kernel_flops": 10 * flops, "kernel_launches": 10}
    r = bench.roofline(info, "fp32")
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s" and r["peak"] == 157.3
    assert abs(r["achieved"] - flops / 150e-6 / 1e12) < 0.01
    assert abs(r["frac"] - r["achieved"] / 157.3) < 1e-3
    assert abs(r["avg_us"] - 150.0) < 1e-6
    json.dumps(r)
    # three real products: the library counts 6 flops per multiply-add and the
    # record keeps them (frac from executed flops); the 8-flop rate is separate
    info3 = dict(info, kernel_flops=info["kernel_flops"] * 0.75)
    r3 = bench.roofline(info3, "fp32", use3m=True)
    assert abs(r3["achieved"] - 0.75 * r["achieved"]) < 0.02
    assert abs(r3["frac"] - r3["achieved"] / 157.3) < 1e-3
    assert abs(r3["effective_tflops_4m"] - r["achieved"]) < 0.02
    assert r3["flops_per_launch"] == 0.75 * flops
    assert bench.roofline({"kernel_launches": 0}, "fp32") is None


def test_cpu_baseline_small():
    rng = np.random.default_rng(0)
    syst = dict(epsilon=1 + rng.random((6, 5, 4)), epsilon_L=1.0, epsilon_R=1.0, k0dx=0.8, xyBC="periodic")
    b = bench.cpu_baseline(syst, n_slices=2)
    assert b["kind"] == "port" and b["unit"] == "slice-updates/s"
    assert b["value"] > 0 and b["cores"] >= 1
    assert "N=30" in b["sample"]
