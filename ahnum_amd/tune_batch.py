"""Batched autotuner: the launch settings of a batch workload, measured on the device.

The reference's `kite tune` times each FFT length and keeps the fastest (tools/kite.c:274-326,
file format src/ctx.c:35-92); this library's `kite tune` does the same per length for one test.
A batch has further settings that no single-test timing sees, and this tool measures them on the
workload itself:

  orders   the shape variant (R x C split, column group width, kernel set) of each FFT length the
           workload uses, i.e. the "order" column of the tuning file ("L us order")
  shape_parts  each FFT shape but the most common on a stream of its own (kite_batch_set_shape_parts)
  parts    streams the tests are dealt to (concurrent launches)
  xcds     XCDs the column jobs are laid out for (1: slot order)
  batch    tests in flight (the rest of a larger workload would wait for a slot); single_large
           only: the other workloads are fixed sets of tests (batch_mixed 256, large_fft 32,
           plain_small 16, record 1), all in flight
  graphs   square calls replayed from a captured HIP graph (true) or each step's kernels launched
           directly on the streams (false)

Each trial is a fresh process (the tuning file is read once per process) that squares the batch
for `--seconds` after a warm-up, in square calls of `--block` steps replayed from the captured
graph (or launched directly), and reports tests/hour. The search is coordinate-wise from the
defaults (orders 0, 2 parts, 8 XCDs, 384 tests, graphs), or from a tuned file's settings
(`--start FILE`), over every setting or those named (`--keys`): orders per length, then shape
parts, parts, XCDs, batch size and graphs, each kept only if it
measures more than `--margin` percent faster (shape parts are tried before the stream count, so that
the count is tuned with them); the winner is then measured twice more against the
defaults (or the start), and kept only if it still wins by the margin. The result is written as JSON (`--out`,
default ahnum_amd/tuned/<workload>.json); bench.py applies the committed file of the workload it
runs (or the one given with `--tuned FILE`).

Usage: python -m ahnum_amd.tune_batch [--workload single_large|batch_mixed|large_fft] [--out FILE]
       [--seconds 3] [--start FILE] [--keys graphs,...]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DEFAULT = {"orders": {}, "shape_parts": 0, "parts": 2, "xcds": 8, "batch": 384}
# workloads that are a fixed set of tests (all in flight): their batch is the set's size, not tuned
FIXED_BATCH = {"batch_mixed": 256, "large_fft": 32, "plain_small": 16, "record": 1}


def defaults(workload):
    d = json.loads(json.dumps(DEFAULT))
    d["batch"] = FIXED_BATCH.get(workload, d["batch"])
    return d


def fft_times_file(orders):
    """A tuning file ("L us order") that sets the shape variant of the given lengths; the time
    column is not used when only some lengths of a range are listed."""
    f = tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False)
    for L, o in sorted(orders.items(), key=lambda kv: int(kv[0])):
        f.write(f"{int(L)} 1.0 {int(o)}\n")
    f.close()
    return f.name


def apply_env(cfg, env=None):
    """Environment for a process that runs with cfg's shape variants (KITE_FFT_TIMES)."""
    env = dict(os.environ if env is None else env)
    if cfg.get("orders"):
        env["KITE_FFT_TIMES"] = fft_times_file(cfg["orders"])
    return env


def trial_main(argv):
    """One trial in this process: tests/hour of the workload's batch under the given settings."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="single_large")
    ap.add_argument("--batch", type=int, default=384)
    ap.add_argument("--parts", type=int, default=2)
    ap.add_argument("--shape-parts", type=int, default=0)
    ap.add_argument("--xcds", type=int, default=8)
    ap.add_argument("--graphs", type=int, default=1)
    ap.add_argument("--block", type=int, default=200)
    ap.add_argument("--seconds", type=float, default=3.0)
    a = ap.parse_args(argv)
    sys.path.insert(0, ROOT)
    from ahnum_amd import kite, workloads
    tests = workloads.record() if a.workload == "record" else getattr(workloads, a.workload)(a.batch)
    b = kite.Batch()
    try:
        b.parts(a.parts)
        b.shape_parts(a.shape_parts)
        b.xcds(a.xcds)
        b.graphs(bool(a.graphs))
        lengths = set()
        for i, (k, n) in enumerate(tests):
            s = b.add(k, n)
            b.set(s, 0, workloads.random_residue_limbs(k * (1 << n) - 1, 100 + i))
            lengths.add(b.fft_len(s))
        b.square(a.block)
        w0 = time.perf_counter()
        while time.perf_counter() - w0 < 0.5:
            b.square(a.block)
            b.sync()
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < a.seconds:
            b.square(a.block)
            b.sync()
            steps += a.block
        dt = time.perf_counter() - t0
        err = max(b.maxerr().values())
    finally:
        b.close()
    tph = sum(steps / dt * 3600.0 / n for _, n in tests)
    print(json.dumps({"tests_per_hour": tph, "lengths": sorted(lengths), "maxerr": err}), flush=True)
    return 0


def run_trial(cfg, workload, seconds, block):
    cmd = [sys.executable, "-m", "ahnum_amd.tune_batch", "--trial", "--workload", workload,
           "--batch", str(cfg["batch"]), "--parts", str(cfg["parts"]), "--xcds", str(cfg["xcds"]),
           "--shape-parts", str(cfg.get("shape_parts", 0)), "--graphs", str(int(cfg.get("graphs", True))),
           "--block", str(block), "--seconds", str(seconds)]
    env = apply_env(cfg)
    try:
        out = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    finally:
        if "KITE_FFT_TIMES" in env and env.get("KITE_FFT_TIMES") != os.environ.get("KITE_FFT_TIMES"):
            os.unlink(env["KITE_FFT_TIMES"])
    if out.returncode != 0:
        raise RuntimeError(f"trial failed ({out.returncode}): {out.stderr[-2000:]}")
    return json.loads(out.stdout.strip().splitlines()[-1])


def merge_refinement(start, result):
    """A search started from a tuned file (--start) keeps that file's record: its trials come first,
    the defaults' rate stays the one it was measured against, and its command is kept beside the
    new one."""
    merged = dict(result)
    merged["trials"] = list(start.get("trials", [])) + list(result["trials"])
    merged["defaults_tests_per_hour"] = start.get("defaults_tests_per_hour", result["defaults_tests_per_hour"])
    merged["commands"] = list(start.get("commands", [start["command"]] if "command" in start else [])) + \
        [result["command"]]
    merged.pop("command", None)
    return merged


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--trial", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--workload", default="single_large")
    ap.add_argument("--out", default=None, help="result file (default: ahnum_amd/tuned/<workload>.json)")
    ap.add_argument("--seconds", type=float, default=3.0)
    ap.add_argument("--block", type=int, default=200)
    ap.add_argument("--variants", type=int, default=4, help="shape variants tried per length")
    ap.add_argument("--batches", default="256,320,384,448")
    ap.add_argument("--start", default=None,
                    help="a tuned file whose best settings the search starts from (default: the defaults)")
    ap.add_argument("--keys", default=None,
                    help="comma-separated settings to search (orders, shape_parts, parts, xcds, batch, graphs; "
                         "default: all)")
    ap.add_argument("--margin", type=float, default=1.0,
                    help="percent a setting must gain to replace the current one (trial noise is ~0.5%%)")
    a, rest = ap.parse_known_args(argv)
    if a.trial:
        return trial_main(sys.argv[2:] if argv is None else argv[1:])
    trials = []

    def measure(cfg, what):
        r = run_trial(cfg, a.workload, a.seconds, a.block)
        trials.append({"tried": what, "config": json.loads(json.dumps(cfg)), **r})
        print(f"{what:28s} {r['tests_per_hour']:10.1f} tests/h  lengths {r['lengths']}  maxerr {r['maxerr']:.4f}",
              flush=True)
        return r

    if a.out is None:
        a.out = os.path.join(HERE, "tuned", a.workload + ".json")
    dflt_cfg = defaults(a.workload)
    start = None
    if a.start:
        with open(a.start) as f:
            start = json.load(f)
        dflt_cfg = start["best"]
    keys = set(a.keys.split(",")) if a.keys else {"orders", "shape_parts", "parts", "xcds", "batch", "graphs"}
    best = json.loads(json.dumps(dflt_cfg))
    base = measure(best, "start" if a.start else "defaults")
    best_tph = base["tests_per_hour"]
    for L in (base["lengths"] if "orders" in keys else []):
        for o in range(1, a.variants):
            cfg = json.loads(json.dumps(best))
            cfg["orders"][str(L)] = o
            r = measure(cfg, f"order L={L} -> {o}")
            if r["tests_per_hour"] > best_tph * (1 + a.margin / 100):
                best, best_tph = cfg, r["tests_per_hour"]
    for key, values in (("shape_parts", [0, 1]), ("parts", [1, 2, 3, 4]), ("xcds", [1, 8]),
                        ("batch", [] if a.workload in FIXED_BATCH else [int(v) for v in a.batches.split(",")]),
                        ("graphs", [True, False])):
        if key not in keys:
            continue
        for v in values:
            if v == best.get(key, True if key == "graphs" else None):
                continue
            cfg = json.loads(json.dumps(best))
            cfg[key] = v
            r = measure(cfg, f"{key} -> {v}")
            if r["tests_per_hour"] > best_tph * (1 + a.margin / 100):
                best, best_tph = cfg, r["tests_per_hour"]
    # the winner against the defaults once more, each measured twice, so that one lucky trial
    # does not decide
    if best != dflt_cfg:
        again = [measure(best, "confirm best")["tests_per_hour"] for _ in range(2)]
        dflt = [measure(json.loads(json.dumps(dflt_cfg)), "confirm defaults")["tests_per_hour"] for _ in range(2)]
        if sum(again) <= sum(dflt) * (1 + a.margin / 100):
            best = json.loads(json.dumps(dflt_cfg))
        best_tph = max(sum(again), sum(dflt)) / 2
    result = {"workload": a.workload, "best": best, "tests_per_hour": best_tph,
              "defaults_tests_per_hour": base["tests_per_hour"], "seconds_per_trial": a.seconds,
              "block": a.block, "trials": trials,
              "command": "python -m ahnum_amd.tune_batch " + " ".join(sys.argv[1:] if argv is None else argv)}
    if start is not None:
        result = merge_refinement(start, result)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(result, f, indent=1)
    print("best", json.dumps(best), f"{best_tph:.1f} tests/h (defaults {base['tests_per_hour']:.1f})", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
