"""CPU checks of what bench.py reports without a GPU: the refusal of more than
one GPU, and the measured-traffic table it reads for `roofline.traffic`."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_table_covers_the_default_shapes():
    """The profiled HBM bytes per launch, keyed by kernel kind and tile shape:
    both two-step shapes the 64-shot workload runs (heuristic 250-row and
    autotuned 500-row segments) and the one-step shape, each between the
    compulsory bytes of one step and twice them, with the copy-kernel
    calibration factors the counters gave."""
    prof = json.load(open(os.path.join(ROOT, "profiles", "psv_2000x1000_traffic.json")))
    compulsory = prof["compulsory_bytes_per_launch"]
    assert compulsory == 64 * 2000 * 1000 * 4 * (10 + 3 / 64)
    for shape in ("512,250,1", "512,500,1"):
        e = prof["two_step"][shape]
        assert compulsory < e["hbm_bytes_per_launch"] < 1.2 * compulsory, shape
        assert abs(e["fetch_factor"] - 2.0) < 1e-3 and abs(e["write_factor"] - 1.0) < 1e-3
        assert e["hbm_bytes_per_launch"] == e["fetch_kib_per_dispatch"] * e["fetch_factor"] * 1024 + \
            e["write_kib_per_dispatch"] * e["write_factor"] * 1024
        assert 0 < e["receiver_kernel"]["hbm_bytes_per_launch"] < 0.02 * compulsory
    one = prof["one_step"]["256,500,4"]
    assert compulsory < one["hbm_bytes_per_launch"] < 1.2 * compulsory
    assert prof["algorithmic_bytes_per_launch"] == {"two_step": 2 * compulsory, "one_step": compulsory}
