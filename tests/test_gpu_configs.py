"""The large-FFT batch and the serving stream of the benchmark configurations, on the GPU.

  large-FFT batch: 32 Fermat base-3 tests, k odd in [3, 999], n in [3M, 6M] (workloads.large_fft),
                   FFT 204800-442368, as one batch through the shared launches (two adjacent column
                   lengths per launch): two squarings of every slot against exact integer squaring,
                   then the production loop (Gerbicz blocks and checks) for a bounded time; and the
                   reference's three smallest large-FFT residues (tests/large_fft_cases.csv) to
                   completion as one batch.
  serving:         the Poisson stream of Fermat and LLR requests (workloads.serve_stream, n
                   log-uniform in [200K, 5M]) through kite_batch_run with a slot limit; every request
                   completed in the mix gives the result it gives alone (kite_prp / kite_llr).
"""
import random
import sys

import pytest

from ahnum_amd import kite, workloads
import bigint

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_large_fft_batch_exact_squarings():
    tests = workloads.large_fft(32)
    r = random.Random(4)
    b = kite.Batch()
    try:
        vals = []
        for k, n in tests:
            s = b.add(k, n)
            a = r.randrange(k * (1 << n) - 1)
            b.set(s, 0, a)
            vals.append((s, k, n, a))
        lengths = sorted({b.fft_len(s) for s, *_ in vals})
        say("large-FFT batch: 32 tests, lengths", lengths)
        b.timing(1)  # the launches of both steps, with the kernels' names
        b.square(2)
        b.sync()
        names = [name for name, *_ in b.kernel_times()]
        say("kernels", names)
        assert any("_rr<" in name for name in names)  # adjacent column lengths share launches
        errs = b.maxerr()
        for s, k, n, a in vals:
            assert b.get(s, 0) == bigint.sqr_n(a, 2, k, n), (k, n)
            assert errs[s] < 0.4
    finally:
        b.close()


def test_large_fft_batch_gerbicz_checked():
    """The 32 tests from x_0 = 3^k with Gerbicz blocks of 500 squarings and a check every 4 blocks,
    under a 15 s horizon: every test passes at least one check, none fails or retries."""
    tests = workloads.large_fft(32)
    rc, out = kite.batch_run([(kite.JOB_PRP, k, n, 0.0) for k, n in tests], block=500, blocks=4, horizon=15.0)
    for (k, n), res in zip(tests, out):
        say(k, n, "status", res.status, "fft", res.fft_len, "verified", res.verified, "iter", res.iter,
            "maxerr", round(res.maxerr, 4))
        assert res.status == kite.JOB_PENDING, (k, n)
        assert res.verified >= 2000 and res.verified % 2000 == 0, (k, n)
        assert res.gerbicz_errors == 0 and res.roundoff_retries == 0 and 0 < res.maxerr < 0.4, (k, n)


def test_large_fft_reference_residues_smallest_three(golden):
    """The three smallest-n rows of the reference's tests/large_fft_cases.csv (n 2.29M-3.32M, FFT
    163840-262144) to completion as one batch: every RES64 equals the reference's."""
    rows = sorted(golden("large_fft_cases.csv"), key=lambda row: int(row[1]))[:3]
    tests = [(int(k), int(n)) for k, n, *_ in rows]
    say("large-FFT residues", tests)
    rc, res = kite.batch_prp(tests)
    assert rc == 0
    for (k, n, want, ref_fft, _src), (status, r) in zip(rows, res):
        say(k, n, status, "%016X" % r.res64, "fft", r.fft_len, "(reference", ref_fft + ")", "maxerr", r.maxerr)
        assert status == 0 and "%016X" % r.res64 == want and r.maxerr < 0.4, (k, n)


def test_serving_stream_alone_equals_mix():
    """The serving stream (64 requests at 20/s, half Fermat half LLR, n log-uniform in [200K, 5M])
    with 16 slots, least work first, rounds shared by device time, for 60 s: requests complete, none
    fails, and each of the (up to) four smallest completed requests gives the same answer alone (the
    LLR requests that complete inside the horizon are compared too; LLR in a mixed batch is also
    covered by tests/test_gpu_batch.py)."""
    stream = workloads.serve_stream(count=64, rate=20.0)
    jobs = [(kite.JOB_LLR if kind == "llr" else kite.JOB_PRP, k, n, t) for t, kind, k, n in stream]
    rc, out = kite.batch_run(jobs, slots=16, order=kite.ORDER_LEAST_WORK, share=kite.SHARE_TIME, horizon=60.0)
    done = [(j, r) for j, r in zip(jobs, out) if r.status == 0]
    say("completed", len(done), "of", len(jobs), "failed",
        sum(1 for r in out if r.status not in (0, kite.JOB_PENDING)))
    assert all(r.status in (0, kite.JOB_PENDING) for r in out)
    tested = [(j, r) for j, r in done if not (j[0] == kite.JOB_LLR and r.factor)]
    assert len(tested) >= 2
    fermat = sorted((jr for jr in tested if jr[0][0] == kite.JOB_PRP), key=lambda jr: jr[0][2])
    llr = sorted((jr for jr in tested if jr[0][0] == kite.JOB_LLR), key=lambda jr: jr[0][2])
    for (kind, k, n, t), r in fermat[:3] + llr[:2]:
        say("alone", "llr" if kind == kite.JOB_LLR else "prp", k, n, "in the mix", "%016X" % r.res64, r.is_prime)
        if kind == kite.JOB_PRP:
            rc1, alone = kite.prp(k, n)
            assert rc1 == 0 and alone.res64 == r.res64 and bool(alone.is_prp) == bool(r.is_prime), (k, n)
        else:
            rc1, alone = kite.llr(k, n)
            assert rc1 == 0 and alone.res64 == r.res64 and bool(alone.is_prime) == bool(r.is_prime), (k, n)
            assert alone.P == r.P
