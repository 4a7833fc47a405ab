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


@pytest.mark.parametrize("name", ["single_large", "batch_mixed", "large_fft", "plain_small"])
def test_tuned_files_fit_their_workload(name):
    """Each committed batched-tuner result names its workload, keeps a fixed workload's test count,
    sets only shape variants the library has (orders 0..3) for lengths the workload uses, chooses
    graph replay or direct launches, and beats the defaults it was measured against."""
    import json
    with open(os.path.join(bench.TUNED_DIR, name + ".json")) as f:
        t = json.load(f)
    assert t["workload"] == name
    best = t["best"]
    assert set(best) - {"graphs"} == {"orders", "shape_parts", "parts", "xcds", "batch"}
    assert 1 <= best["parts"] <= 4 and best["xcds"] in (1, 8) and best["shape_parts"] in (0, 1)
    assert best.get("graphs", True) in (True, False)
    wcount = bench.WORKLOADS[name][2]
    if wcount is not None:
        assert best["batch"] == wcount
    lengths = {L for tr in t["trials"] for L in tr["lengths"]}
    for L, o in best["orders"].items():
        assert int(L) in lengths and 1 <= int(o) <= 3
    assert t["tests_per_hour"] > t["defaults_tests_per_hour"]


@pytest.mark.parametrize("workload,kernel", [
    ("single_large", "k_cols_fused_r<96>"), ("record", "k_cols_fused_r<400>"), ("record", "k_rows_c<2048>"),
])
def test_traffic_summaries_give_the_dominant_kernels(workload, kernel):
    """bench.py reads HBM bytes per launch from the committed rocprofv3 summaries: at least the
    kernel's algorithmic bytes (16 L per test in the launch: the residue read and written; 307
    tests at L = 49152 in the n~660K batch's main column launch) and at most 1.5 times them."""
    b = bench.load_traffic(workload, kernel)
    L, tests = {"single_large": (49152, 307), "record": (1638400, 1)}[workload]
    alg = 16.0 * L * tests
    assert b is not None and 1.0 <= b / alg <= 1.5
    assert bench.load_traffic("plain_small", kernel) is None


def test_tuner_refinement_keeps_the_start_record():
    """A search started from a tuned file (--start) keeps that file's trials first, the rate of the
    defaults it was measured against and its command beside the new one."""
    from ahnum_amd import tune_batch
    start = {"workload": "w", "best": {"parts": 2}, "tests_per_hour": 10.0, "defaults_tests_per_hour": 8.0,
             "trials": [{"tried": "defaults"}, {"tried": "parts -> 2"}], "command": "python -m ahnum_amd.tune_batch"}
    new = {"workload": "w", "best": {"parts": 2, "graphs": False}, "tests_per_hour": 11.0,
           "defaults_tests_per_hour": 10.0, "trials": [{"tried": "start"}, {"tried": "graphs -> False"}],
           "command": "python -m ahnum_amd.tune_batch --start f --keys graphs"}
    m = tune_batch.merge_refinement(start, new)
    assert [t["tried"] for t in m["trials"]] == ["defaults", "parts -> 2", "start", "graphs -> False"]
    assert m["defaults_tests_per_hour"] == 8.0 and m["tests_per_hour"] == 11.0
    assert m["commands"] == [start["command"], new["command"]] and "command" not in m
    again = tune_batch.merge_refinement(m, dict(new, command="second"))
    assert again["commands"] == [start["command"], new["command"], "second"]
