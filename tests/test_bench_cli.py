"""bench.py's contract without a GPU: it refuses more than one GPU with a message, and each batch
workload draws the tests its configuration names (counts, k odd in range, n in range, seeded so
that repeated runs measure the same numbers)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from ahnum_amd import workloads  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr


@pytest.mark.parametrize("name,count,k_range,n_range", [
    ("plain_small", 16, (1001, 9999), (60000, 70000)),
    ("batch_mixed", 256, (3, 9999), (300000, 700000)),
    ("large_fft", 32, (3, 999), (3000000, 6000000)),
    ("record", 1, (3, 99), (24000000, 25000000)),
])
def test_workloads_draw_their_configuration(name, count, k_range, n_range):
    metric, described, wcount, seed = bench.WORKLOADS[name]
    tests = workloads.record() if name == "record" else getattr(workloads, name)(wcount)
    assert len(tests) == count == wcount
    for k, n in tests:
        assert k % 2 == 1 and k_range[0] <= k <= k_range[1]
        assert n_range[0] <= n <= n_range[1]
    assert tests == (workloads.record(seed=seed) if name == "record" else getattr(workloads, name)(wcount, seed=seed))
    assert len(set(tests)) == len(tests)


def test_default_workload_is_the_metric_batch():
    metric, described, count, seed = bench.WORKLOADS["single_large"]
    assert count is None  # sized by --batch (384 by default)
    tests = workloads.single_large(384, seed=seed)
    assert all(k % 2 == 1 and 1001 <= k <= 9999 and 650000 <= n <= 670000 for k, n in tests)
    assert "n~660K" in metric
