"""bench.py's command line: more than one GPU is refused with a message (the
build runs one replica chain on one GPU; independent replicas would shard
across GPUs without communication), and the traffic record it reports is read
from the committed counter-pass summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "--gpus 2 is not supported" in r.stderr and "one GPU" in r.stderr


def test_traffic_record_is_read():
    sys.path.insert(0, ROOT)
    import bench
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        rec = json.load(f)["k_da_chain"]
    assert bench._traffic("k_da_chain") == rec["bytes_per_launch"] > 0
    assert bench._traffic("no_such_kernel") is None
