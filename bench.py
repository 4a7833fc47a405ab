#!/usr/bin/env python3
"""Benchmark of the MI355X squaring engine on the Fermat workload at n ~ 660K (FFT ~49152).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--cpu-secs S] [--no-cpu]
                  [--gerbicz-secs S] [--serve-secs S]
  python bench.py --workload {batch_mixed,large_fft,record,plain_small} [...]
  python bench.py --serve [--serve-count C] [--serve-rate R] [--serve-horizon S] ...

A step is one squaring of every residue in the batch (B Fermat base-3 tests with k odd in
[1001, 9999], n in [650000, 670000], seeded; B = 384 by default): the batched engine advances all
B tests with one launch per pass and stream. `value` is the squaring throughput of the K timed
steps as tests/hour, the reference's throughput formula (tools/kite.c:262): sum over tests of
squarings/s * 3600 / n. The K steps run in square calls of min(K, 1000) steps (a Gerbicz block),
replayed from the HIP graph the batch captured for that call; the graph is captured and replayed
once in full before the timed region, then W warm-up steps run.

Beside it, one JSON line carries:
  plain          the same number alone through kite.h (10^4 squarings of a seeded random residue):
                 us/squaring and max round-off, the one-test-at-a-time GPU path
  roofline       the dominant pass kernel (the most device time) in a second region of launches
                 timed with HIP events on the batch stream: algorithmic bytes per launch (16 L per
                 test: each pass reads and writes the 8L-byte residue) / average launch duration
  path_roofline  the whole step against HBM (32 L bytes per test) and FP64 (5 L log2(L/2) + 20 L)
  gerbicz_checked  the same tests through kite_batch_run, the production loop (Gerbicz blocks of
                 1000 with their multiplies, a check every 64 blocks on the device, round-off reads)
                 for a bounded time: its squarings/s as tests/hour, and the verified share
  serve          a bounded serving run: 200 Poisson arrivals (20/s) of Fermat and LLR requests (n
                 log-uniform in [200K, 5M]), at most 16 in flight, least work first: p50/p99
                 completion latency per kind, backlog at the horizon
  cpu_baseline   the oracle's restatement of the reference's portable two-pass backend (c2p),
                 one test per host core, timed on this host
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# the batch workloads (ahnum_amd/workloads.py): metric, description, test count (None: --batch), seed
WORKLOADS = {
    "single_large": ("tests/hour at n~660K (Fermat base 3, FFT 49152, Gerbicz-checkable residues), batched",
                     "single_large batch: B Fermat base-3 tests, k odd in [1001,9999], n in [650000,670000]",
                     None, 2),
    "batch_mixed": ("tests/hour, 256 Fermat base-3 tests at n 300K-700K (mixed FFT lengths), batched",
                    "batch_mixed: 256 Fermat base-3 tests, k odd in [3,9999], n in [300000,700000]", 256, 3),
    "large_fft": ("tests/hour, 32 Fermat base-3 tests at n 3M-6M (FFT 160K-442K), batched",
                  "large_fft: 32 Fermat base-3 tests, k odd in [3,999], n in [3000000,6000000]", 32, 4),
    "record": ("tests/hour of a record-size Fermat test at n 24M-25M (FFT ~1.6M)",
               "record: 1 Fermat base-3 test, k odd < 100, n in [24000000,25000000]", 1, 5),
    "plain_small": ("tests/hour, 16 Fermat base-3 tests at n 60K-70K (FFT ~4K), batched",
                    "plain_small: 16 Fermat base-3 tests, k odd in [1001,9999], n in [60000,70000]", 16, 1),
}

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (FMA counted as two)


# ---- CPU baseline ------------------------------------------------------------------------------

def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_has_v3():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return {"avx2", "fma", "bmi2"} <= flags
    except OSError:
        pass
    return False


def host_cores():
    """Cores this process may use: its CPU affinity, within the share the host grants the job
    (OMP_NUM_THREADS, when the host sets it to announce that share)."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(share)) if share.isdigit() and int(share) > 0 else n


def cpu_worker(args):
    k, n, secs = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    from ahnum_amd import workloads
    c = O.OCtx(k, n, 0, O.OCtx.C2P)
    x = c.num(workloads.random_residue_limbs(c.N, 2))
    O.lib().or_square(c.p, x)
    t0 = time.time()
    it = 0
    while time.time() - t0 < secs:
        O.lib().or_square(c.p, x)
        it += 1
    dt = time.time() - t0
    L = c.L
    c.close()
    return it / dt, L


def cpu_baseline(tests, secs):
    cores = host_cores()
    fast = os.path.join(ROOT, "oracle", "bin", "liboracle_v3.so")
    isa = "x86-64-v2, -O2"
    if cpu_has_v3() and os.path.exists(fast):
        os.environ["KITE_ORACLE_LIB"] = fast  # inherited by the spawned workers
        isa = "x86-64-v3, -O3"
    jobs = [(k, n, secs) for k, n in (tests * cores)[:cores]]
    with mp.get_context("spawn").Pool(cores) as pool:
        out = pool.map(cpu_worker, jobs)
    os.environ.pop("KITE_ORACLE_LIB", None)
    tph = sum(rate * 3600.0 / n for (rate, _), (_, n, _) in zip(out, jobs))
    us = 1e6 / (sum(r for r, _ in out) / len(out))
    return {
        "value": round(tph, 3),
        "unit": "tests/hour",
        "cores": cores,
        "kind": "port",
        "cpu": cpu_model(),
        "build": isa,
        "sample": f"c2p restatement (oracle/kite_oracle.c, {isa}), one test per core on {cores} cores, "
                  f"{secs:.0f} s of squarings per core at L={out[0][1]} ({us:.0f} us/squaring/core), "
                  f"extrapolated as rate*3600/n",
    }


# ---- traffic summary ---------------------------------------------------------------------------

# the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of each workload's benchmark command
TRAFFIC_FILES = {w: os.path.join(ROOT, "profiles", f"hbm_traffic_{f}.txt")
                 for w, f in (("single_large", "single_large_batch"), ("large_fft", "large_fft_batch"),
                              ("record", "record"))}
TUNED_DIR = os.path.join(ROOT, "ahnum_amd", "tuned")  # <workload>.json: the batched autotuner's result


def load_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    of the workload's benchmark (lines "traffic <kernel> <bytes per launch>", given for kernels that
    run one launch shape), or None."""
    if workload not in TRAFFIC_FILES:
        return None
    try:
        with open(TRAFFIC_FILES[workload]) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic" and w[1] == kernel:
                    return float(w[2])
    except (OSError, ValueError):
        pass
    return None


# ---- serving -----------------------------------------------------------------------------------

def pct(values, p):
    v = sorted(values)
    return round(v[min(len(v) - 1, int(p * len(v)))], 3) if v else None


def run_serve(count, rate, n_lo, n_hi, horizon, hbm_gb, slots, order_name, ckpt_dir="", ckpt_secs=0.0,
              verbose=False, share_name="squarings", launch_name="graph"):
    """A finite Poisson stream of Fermat and LLR requests through kite_batch_run; the serving record."""
    from ahnum_amd import kite, workloads
    stream = workloads.serve_stream(count=count, rate=rate, n_lo=n_lo, n_hi=n_hi)
    jobs = [(kite.JOB_LLR if kind == "llr" else kite.JOB_PRP, k, n, t) for t, kind, k, n in stream]

    def progress(i, r):  # one line per completed request, on stderr
        if verbose:
            print(f"done {i} status {r.status} n {jobs[i][2]} latency {r.finish - jobs[i][3]:.2f}s",
                  file=sys.stderr, flush=True)

    if ckpt_dir:
        os.makedirs(ckpt_dir, exist_ok=True)
    order = kite.ORDER_LEAST_WORK if order_name == "least_work" else kite.ORDER_ARRIVAL
    t0 = time.perf_counter()
    rc, out = kite.batch_run(jobs, slots=slots, hbm_budget=int(hbm_gb * 2**30), block=1000, horizon=horizon,
                             on_done=progress, checkpoint_dir=ckpt_dir or None, checkpoint_secs=ckpt_secs,
                             order=order, share=kite.SHARE_TIME if share_name == "time" else kite.SHARE_SQUARINGS,
                             launch=kite.LAUNCH_DIRECT if launch_name == "direct" else kite.LAUNCH_GRAPH)
    wall = time.perf_counter() - t0
    span = min(wall, horizon) if horizon else wall
    arrived = sum(1 for j in jobs if j[3] <= span)
    factor = [(j, r) for j, r in zip(jobs, out) if r.status == 0 and j[0] == kite.JOB_LLR and r.factor]
    tested = [(j, r) for j, r in zip(jobs, out) if r.status == 0 and not (j[0] == kite.JOB_LLR and r.factor)]
    lat = {"fermat": [r.finish - j[3] for j, r in tested if j[0] == kite.JOB_PRP],
           "llr": [r.finish - j[3] for j, r in tested if j[0] == kite.JOB_LLR]}
    rec = {
        "requests": count, "rate_per_s": rate, "n_range": [n_lo, n_hi], "horizon_s": horizon,
        "hbm_budget_gib": hbm_gb, "slots": slots, "order": order_name, "share": share_name, "block": 1000,
        "launch": launch_name,
        "arrived": arrived,
        "completed": {"fermat": len(lat["fermat"]), "llr": len(lat["llr"])},
        "completed_tests_per_hour": round(len(tested) * 3600.0 / span, 1),
        "llr_factor_found": len(factor),
        "latency_p50_s": {kind: pct(v, 0.50) for kind, v in lat.items()},
        "latency_p99_s": {kind: pct(v, 0.99) for kind, v in lat.items()},
        "max_completed_n": max((j[2] for j, _ in tested), default=0),
        "in_flight_at_horizon": sum(1 for r in out if r.status == kite.JOB_PENDING and r.fft_len),
        "waiting_at_horizon": sum(1 for j, r in zip(jobs, out) if r.status == kite.JOB_PENDING and not r.fft_len
                                  and j[3] <= span),
        "failed": sum(1 for r in out if r.status not in (0, kite.JOB_PENDING)),
        "maxerr": round(max((r.maxerr for _, r in tested), default=0.0), 4),
        "wall_s": round(wall, 2),
    }
    if ckpt_dir:  # restartable serving: the stopped jobs' states are on disk
        rec["checkpoint_dir"] = ckpt_dir
        rec["resumed"] = sum(1 for r in out if r.resumed)
        rec["checkpoint_files"] = sum(1 for f in os.listdir(ckpt_dir) if f.endswith(".ckpt"))
    return rec


def serve_only(a):
    rec = run_serve(a.serve_count, a.serve_rate, a.serve_n_lo, a.serve_n_hi, a.serve_horizon, a.serve_hbm_gb,
                    a.serve_slots, a.serve_order, a.serve_checkpoint_dir, a.serve_checkpoint_secs, verbose=True,
                    share_name=a.serve_share, launch_name=a.serve_launch)
    line = {
        "metric": "serving: completed tests/hour (Poisson arrivals, 50% Fermat + 50% LLR, n log-uniform)",
        "value": rec["completed_tests_per_hour"],
        "unit": "tests/hour",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic request stream (seeded)",
        "config": {"workload": "serve_stream"},
        "serve": rec,
    }
    print(json.dumps(line), flush=True)


# ---- the Gerbicz-checked production loop ---------------------------------------------------------

def gerbicz_leg(tests, secs, blocks, launch):
    """The tests through kite_batch_run for `secs`: Fermat from x_0 = 3^k, Gerbicz blocks of 1000
    squarings with their multiplies, a check every `blocks` blocks, round-off read every round.
    Rates start once every test has been admitted (their starts, x_0 = 3^k by powering, come
    first): tests_per_hour counts every squaring the loop did (those after a test's last check are
    covered by its next one), verified_tests_per_hour only those up to each test's last passed
    check (at a horizon of a few checks that undercounts by up to one check interval)."""
    from ahnum_amd import kite
    jobs = [(kite.JOB_PRP, k, n, 0.0) for k, n in tests]
    t0 = time.perf_counter()
    rc, out = kite.batch_run(jobs, block=1000, blocks=blocks, horizon=secs, **launch)
    wall = time.perf_counter() - t0
    admitted = max(r.start for r in out)  # the last test's setup done: all are in the rounds
    span = max(min(wall, secs) - admitted, 1e-9)
    verified = sum(r.verified for r in out)
    done = sum(r.iter for r in out)
    tph = sum(r.iter / span * 3600.0 / n for r, (_, n) in zip(out, tests))
    vtph = sum(r.verified / span * 3600.0 / n for r, (_, n) in zip(out, tests))
    return {
        "tests_per_hour": round(tph, 1),
        "verified_tests_per_hour": round(vtph, 1),
        "gerbicz_block": 1000,
        "blocks_per_check": blocks,
        "setup_seconds": round(admitted, 2),
        "seconds": round(span, 2),
        "verified_squarings": verified,
        "squarings": done,
        "gerbicz_errors": sum(r.gerbicz_errors for r in out),
        "roundoff_retries": sum(r.roundoff_retries for r in out),
        "maxerr": round(max(r.maxerr for r in out), 4),
        "failed": sum(1 for r in out if r.status not in (0, kite.JOB_PENDING)),
    }


# ---- main --------------------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--batch", type=int, default=None,
                    help="tests of the single_large workload (default: the tuned value, else 384)")
    ap.add_argument("--tuned", default="auto",
                    help="batched autotuner result (python -m ahnum_amd.tune_batch) applied to the workload's "
                         "batch: shape variant per length, shape parts, streams, XCD layout (and the batch size "
                         "of single_large); 'auto' takes ahnum_amd/tuned/<workload>.json when present, 'none' "
                         "runs the defaults")
    ap.add_argument("--workload", default="single_large", choices=sorted(WORKLOADS),
                    help="batch workload (single_large: the n~660K batch the metric is quoted on)")
    ap.add_argument("--plain-steps", type=int, default=10000, help="squarings of the one-test microbench")
    ap.add_argument("--cpu-secs", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gerbicz-secs", type=float, default=30.0,
                    help="length of the Gerbicz-checked leg (0: skip)")
    ap.add_argument("--gerbicz-launch", default="graph", choices=["graph", "direct"],
                    help="rounds of the Gerbicz-checked loop replayed from captured graphs or launched directly "
                         "(graph: measured faster for this loop; the squaring region follows the tuned file)")
    ap.add_argument("--gerbicz-blocks", type=int, default=64,
                    help="blocks of 1000 squarings between checks in the Gerbicz-checked leg (a check "
                         "costs one block: 64 keeps it at 1.6%%)")
    ap.add_argument("--serve-secs", type=float, default=30.0, help="horizon of the serving leg (0: skip)")
    ap.add_argument("--serve-leg-count", type=int, default=200, help="requests of the serving leg")
    ap.add_argument("--serve-leg-slots", type=int, default=16, help="jobs in flight at most in the serving leg")
    ap.add_argument("--time-every", type=int, default=10,
                    help="time the launches of every N-th step of the kernel-timing region with HIP events stamped from their dispatch packets")
    ap.add_argument("--time-steps", type=int, default=1000, help="steps of the kernel-timing region")
    ap.add_argument("--block", type=int, default=1000, help="squarings per square call (a Gerbicz block)")
    ap.add_argument("--parts", type=int, default=None,
                    help="concurrent streams the batch is dealt to (default: the tuned value, else 2)")
    ap.add_argument("--shape-parts", type=int, default=None,
                    help="each FFT shape but the most common on a stream of its own (default: the tuned value, "
                         "else 0)")
    ap.add_argument("--launch", default=None, choices=["graph", "direct"],
                    help="square calls replayed from a captured HIP graph or launched directly "
                         "(default: the tuned file's choice, else graph)")
    ap.add_argument("--xcds", type=int, default=None,
                    help="XCDs the column jobs are laid out for (default: the tuned value, else 8)")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run a fill, a copy and an add over this many MiB (calibrates the HBM "
                         "byte counters of a rocprofv3 --pmc run of this command)")
    ap.add_argument("--serve", action="store_true", help="serving run only (the options below)")
    ap.add_argument("--serve-count", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-n-lo", type=int, default=200000)
    ap.add_argument("--serve-n-hi", type=int, default=5000000)
    ap.add_argument("--serve-horizon", type=float, default=120.0)
    ap.add_argument("--serve-hbm-gb", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=0, help="jobs in flight at most (0 = HBM budget only)")
    ap.add_argument("--serve-order", default="arrival", choices=["arrival", "least_work"],
                    help="admission order among arrived requests")
    ap.add_argument("--serve-share", default="time", choices=["squarings", "time"],
                    help="how jobs in flight share the rounds (kite_run_opts.share)")
    ap.add_argument("--serve-launch", default="direct", choices=["graph", "direct"],
                    help="rounds of the serving runs replayed from captured graphs or launched directly "
                         "(direct: what the batched tuner chose for every tuned workload)")
    ap.add_argument("--serve-checkpoint-dir", default="",
                    help="jobs resume from and write checkpoints here; the horizon saves every job in flight")
    ap.add_argument("--serve-checkpoint-secs", type=float, default=0.0, help="checkpoint interval (0 = 600 s)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (independent tests shard "
              f"across GPUs with no exchange; run one process per GPU)", file=sys.stderr)
        sys.exit(2)

    if a.serve:
        serve_only(a)
        return

    # the batched autotuner's settings; the shape variants reach the library through its tuning
    # file, which it reads once, so they are set before it loads
    tuned_path, tuned = None, {}
    if a.tuned != "none":
        tuned_path = os.path.join(TUNED_DIR, a.workload + ".json") if a.tuned == "auto" else a.tuned
        if os.path.exists(tuned_path):
            with open(tuned_path) as f:
                tuned = json.load(f)["best"]
            if tuned.get("orders"):
                from ahnum_amd import tune_batch
                os.environ["KITE_FFT_TIMES"] = tune_batch.fft_times_file(tuned["orders"])
        else:
            tuned_path = None
    if a.batch is None:
        a.batch = int(tuned.get("batch", 384))
    if a.parts is None:
        a.parts = int(tuned.get("parts", 2))
    if a.xcds is None:
        a.xcds = int(tuned.get("xcds", 8))
    if a.launch is None:
        a.launch = "graph" if tuned.get("graphs", True) else "direct"

    from ahnum_amd import kite, workloads

    if a.copy_calib_mb:
        import torch
        src = torch.empty(a.copy_calib_mb << 17, dtype=torch.float64, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)  # reads and writes copy_calib_mb MiB
            torch.add(src, 1.0, out=dst)  # elementwise kernel: same bytes
        torch.cuda.synchronize()
        del src, dst

    metric, described, count, seed = WORKLOADS[a.workload]
    if count is None:
        tests = workloads.single_large(a.batch)
    elif a.workload == "record":
        tests = workloads.record()
    else:
        tests = getattr(workloads, a.workload)(count)

    # ---- plain path: one test at a time through kite.h ----
    k0, n0 = tests[0]
    with kite.Ctx(k0, n0) as c:
        L0 = c.L
        x = c.num(workloads.random_residue_limbs(c.N, 2))
        x.square(max(a.warmup, 64))
        c.maxerr()
        c.reset_maxerr()
        t0 = time.perf_counter()
        x.square(a.plain_steps)
        plain_err = c.maxerr()  # waits for the squarings
        plain_s = time.perf_counter() - t0
        x.free()
    plain_us = plain_s * 1e6 / a.plain_steps
    plain_tph = 3600.0 / (n0 * plain_us * 1e-6)

    # ---- batched engine ----
    b = kite.Batch()
    parts = b.parts(a.parts)
    shape_parts = b.shape_parts(int(tuned.get("shape_parts", 0)) if a.shape_parts is None else a.shape_parts)
    b.xcds(a.xcds)
    b.graphs(a.launch == "graph")
    Ls = []
    for i, (k, n) in enumerate(tests):
        s = b.add(k, n)
        b.set(s, 0, workloads.random_residue_limbs(k * (1 << n) - 1, 100 + i))
        Ls.append(b.fft_len(s))
    # timed region 1 (the value): square calls of `block` steps, as the Gerbicz loop makes them,
    # replayed from the captured HIP graph or launched directly (--launch). Before it: the capture
    # (with graphs), one full warm call of the same size, and calls until at least W steps and a
    # quarter second of squaring have run (the device's clocks come up under load; a short W alone
    # would time a cold device)
    block = min(a.block, a.steps)
    b.square(block)  # captures the graph, if any (and runs it)
    warm, w0 = block, time.perf_counter()
    while warm < max(a.warmup, 2 * block) or time.perf_counter() - w0 < 0.25:
        b.square(block)
        b.sync()
        warm += block
    b.maxerr(reset=True)
    b.sync()
    t0 = time.perf_counter()
    done = 0
    while done < a.steps:
        c = min(block, a.steps - done)
        b.square(c)
        done += c
    b.sync()
    wall = time.perf_counter() - t0
    errs = b.maxerr()
    # timed region 2 (the roofline): the same squarings launched directly, the launches of every
    # `time_every`-th step timed by HIP events stamped from their own dispatch packets (hipExtLaunchKernel)
    b.parts(1)  # kernels alone, so that each launch's duration is its own
    b.timing(a.time_every)
    b.square(a.time_steps)
    b.sync()
    ms, launches = b.pass_times()
    kstats = b.kernel_times()
    b.close()

    rate = a.steps / wall  # squarings/s per test
    tph = sum(rate * 3600.0 / n for _, n in tests)
    names = ["rows", "cols_inv", "cols_fwd"]
    # dominant kernel: the most device time over the timed region (HIP events on the batch stream)
    kname, kms, klaunch, kbytes = max(kstats, key=lambda t: t[1])
    avg_s = kms / 1e3 / klaunch
    alg_bytes = kbytes / klaunch
    achieved = alg_bytes / avg_s / 1e9
    traffic = load_traffic(a.workload, kname)
    step_bytes = sum(32 * L for L in Ls)
    step_flops = sum(5 * L * math.log2(L / 2) + 20 * L for L in Ls)

    out = {
        "metric": metric,
        "value": round(tph, 3 if tph < 100 else 1),
        "unit": "tests/hour",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "warmup_steps_run": warm,
        "ms_per_step": round(wall * 1e3 / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random residues mod k*2^n-1)",
        "config": {
            "workload": described,
            "batch": len(tests),
            "fft_lengths": sorted(set(Ls)),
            "launch": f"square calls of {block} steps "
                      + ("replayed from a captured HIP graph" if a.launch == "graph" else "launched directly")
                      + f", {parts} streams"
                      + (" for the most common FFT shape and one for each other" if shape_parts else "")
                      + f", column jobs laid out for {a.xcds} XCDs",
            "tuned": ({"file": os.path.relpath(tuned_path, ROOT), **tuned} if tuned_path else None),
            "value_is": "squaring throughput of the timed steps as tests/hour (sum of squarings/s * 3600 / n); "
                        "the production loop with Gerbicz checks is timed apart in gerbicz_checked",
            "seed": seed,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": int(alg_bytes),
            "avg_launch_us": round(avg_s * 1e6, 2),
            "pass_ms": {names[p]: round(ms[p], 3) for p in range(3)},
            "kernels": {n: {"avg_us": round(t / c * 1e3, 2), "launches": c,
                            "gbs": round(by / (t / 1e3) / 1e9, 1)} for n, t, c, by in kstats},
        },
        "plain": {
            "us_per_squaring": round(plain_us, 3),
            "tests_per_hour": round(plain_tph, 2),
            "fft_len": L0,
            "maxerr": round(plain_err, 4),
            "number": f"{k0}*2^{n0}-1",
        },
        "path_roofline": {
            "bytes_per_step": int(step_bytes),
            "flops_per_step": int(step_flops),
            "mem_achieved_gbs": round(step_bytes * a.steps / wall / 1e9, 1),
            "mem_frac": round(step_bytes * a.steps / wall / 1e9 / HBM_PEAK_GBS, 4),
            "fp64_achieved_tflops": round(step_flops * a.steps / wall / 1e12, 2),
            "fp64_frac": round(step_flops * a.steps / wall / 1e12 / FP64_PEAK_TFLOPS, 4),
        },
        "batched_us_per_test_squaring": round(wall * 1e6 / a.steps / len(tests), 4),
        "batched_vs_plain": round(tph / plain_tph, 2),
        "maxerr": round(max(errs.values()), 4),
        "gpu_ms_per_timed_step": round(sum(ms) / max(launches[0], 1), 4),
    }
    if a.gerbicz_secs > 0:
        g = gerbicz_leg(tests, a.gerbicz_secs, a.gerbicz_blocks,
                        {"parts": a.parts, "shape_parts": shape_parts, "xcds": a.xcds,
                         "launch": 0 if a.gerbicz_launch == "graph" else 1})
        g["launch"] = a.gerbicz_launch
        g["vs_value"] = round(g["tests_per_hour"] / out["value"], 4) if out["value"] else None
        out["gerbicz_checked"] = g
    if a.serve_secs > 0 and a.workload == "single_large":
        out["serve"] = run_serve(count=a.serve_leg_count, rate=a.serve_rate, n_lo=a.serve_n_lo,
                                 n_hi=a.serve_n_hi, horizon=a.serve_secs, hbm_gb=a.serve_hbm_gb,
                                 slots=a.serve_leg_slots, order_name="least_work", share_name="time",
                                 launch_name=a.serve_launch)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tests[:16], a.cpu_secs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
