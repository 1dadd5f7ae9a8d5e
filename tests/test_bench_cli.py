"""bench.py's contract on the CPU: more than one GPU is refused with a
message, and the roofline traffic comes from the committed PMC summary."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_stage_traffic_sums_the_stage_kernels():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    t = json.load(open(path))
    assert bench.measured_traffic("k_gemm_f64_dma") == t["k_gemm_f64_dma"]
    assert bench.stage_traffic(("k_gemm_f64_dma", "k_ksplit_sum")) == \
        t["k_gemm_f64_dma"] + t.get("k_ksplit_sum", 0)
    assert bench.stage_traffic(("no_such_kernel", "k_ksplit_sum")) is None
