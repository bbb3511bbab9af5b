"""bench.py without a GPU: more than one GPU is refused with a JSON message and exit status 2
before any device work, and the CPU-baseline host record is filled in."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2, r.stderr
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_host_record():
    sys.path.insert(0, ROOT)
    import bench
    h = bench._host_cpu()
    assert h["logical_cpus"] >= 1 and 1 <= h["usable_cpus"] <= h["logical_cpus"]
    assert h["cpu_model"] is None or isinstance(h["cpu_model"], str)
