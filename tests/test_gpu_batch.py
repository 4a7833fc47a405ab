"""The batched engine: many tests of different k, n and FFT length advance through shared launches.
Each result must equal the exact value and the test's result alone."""
import os
import random
import sys

import pytest

from ahnum_amd import kite, workloads
from test_gpu_engine import mod_riesel, sqr_n

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_mixed_lengths_one_launch():
    """Slots of different lengths (and shapes) squared by the same launches, with different
    squaring counts per slot; every value exact."""
    cases = [(4523, 67231), (9357, 20000), (1, 44497), (3, 26459), (8045, 659321), (9999, 30011), (7, 4001)]
    b = kite.Batch()
    r = random.Random(3)
    vals = {}
    try:
        for k, n in cases:
            s = b.add(k, n)
            a = r.randrange(k * (1 << n) - 1)
            b.set(s, 0, a)
            vals[s] = (k, n, a)
        say("lengths", sorted({b.fft_len(s) for s in vals}))
        cap = b.lib.kite_batch_capacity(b.p)
        import ctypes
        which = (ctypes.c_int32 * cap)(*[0] * cap)
        counts = [5 + 3 * s for s in range(cap)]
        cnt = (ctypes.c_uint32 * cap)(*counts)
        assert b.lib.kite_batch_square(b.p, which, cnt) == 0
        for s, (k, n, a) in vals.items():
            assert b.get(s, 0) == sqr_n(a, counts[s], k, n), (k, n)
        errs = b.maxerr()
        assert max(errs.values()) < 0.4
    finally:
        b.close()


@pytest.mark.parametrize("cases", [
    # the n~660K boundary: L 49152 (R 96) and 51200 (R 100) share one launch per column pass
    [(8045, 659321, 0), (9001, 669990, 0), (3, 658000, 0), (9999, 669000, 0)],
    # small adjacent column lengths from explicit FFT lengths
    [(9357, 20000, 0), (4523, 20011, 2560), (3, 20000, 3072), (7, 19000, 3584)],
])
def test_adjacent_lengths_share_column_launch(cases):
    """Tests whose column lengths are adjacent compile-time sizes run through the two-shape column
    kernels; every value exact after several squarings."""
    b = kite.Batch()
    r = random.Random(11)
    vals = {}
    try:
        for k, n, L in cases:
            s = b.add(k, n, fft_len=L)
            a = r.randrange(k * (1 << n) - 1)
            b.set(s, 0, a)
            vals[s] = (k, n, a)
        say("lengths", sorted({b.fft_len(s) for s in vals}))
        b.square(6)
        for s, (k, n, a) in vals.items():
            assert b.get(s, 0) == sqr_n(a, 6, k, n), (k, n, b.fft_len(s))
        assert max(b.maxerr().values()) < 0.4
    finally:
        b.close()


def test_batch_prp_continuous_admission_vs_gmp(golden):
    """The 16 seeded n~65K tests through 5 slots (tests admitted as others finish); RES64s equal
    GMP's, i.e. the same as each test alone."""
    rows = golden("res64_plain_small.csv")
    tests = [(int(k), int(n)) for k, n, _, _ in rows]
    done = []
    rc, out = kite.batch_prp(tests, slots=5, block=1000, on_done=lambda i, st, r: done.append(i))
    assert rc == 0
    assert sorted(done) == list(range(16))
    for (k, n, base, want), (st, r) in zip(rows, out):
        assert st == 0
        assert "%016X" % r.res64 == want, (k, n)
        assert r.maxerr < 0.4


def test_batch_prp_fault_caught():
    """A +1 digit corruption in one test of the batch is caught by its Gerbicz check; every
    result stays right (CMakeLists.txt:95's case, plus neighbours)."""
    tests = [(1203, 90001), (9357, 20000), (8191, 25000)]
    rc, out = kite.batch_prp(tests, block=1000, fault_test=0, fault_at=50000)
    assert rc == 0
    assert "%016X" % out[0][1].res64 == "4BE9F1953E350085"
    assert out[0][1].gerbicz_errors >= 1
    assert "%016X" % out[1][1].res64 == "79291CD2F47CE48C"
    assert "%016X" % out[2][1].res64 == "D45538E6F07266C6"
    assert out[1][1].gerbicz_errors == 0


def test_batch_prp_primecrunch_rows(golden):
    """The reference's 14 primecrunch residues near 9357*2^369717-1 (tests/known_residues.csv:14-27),
    all in one batch."""
    rows = [r for r in golden("known_residues.csv") if r[3] == "primecrunch"]
    assert len(rows) == 14
    tests = [(int(k), int(n)) for k, n, _, _ in rows]
    say("batch of", len(tests), "at n~370K")
    rc, out = kite.batch_prp(tests, block=1000)
    assert rc == 0
    for (k, n, want, _), (st, r) in zip(rows, out):
        assert st == 0 and "%016X" % r.res64 == want, (k, n)
        assert r.maxerr < 0.4


def test_batch_reference_vectors_at_full_size(golden):
    """The reference's own full-size vectors: 9357*2^650001-1 (tools/kite-bench.sh:28-29) and the
    prime 9001*2^662695-1 (tests/benchmarks.csv:2), at FFT length 49152, in one batch."""
    rows = [r for r in golden("single_vectors.csv") if int(r[1]) >= 600000]
    assert len(rows) == 2
    tests = [(int(k), int(n)) for k, n, _, _ in rows]
    say("batch of", tests)
    rc, out = kite.batch_prp(tests, block=1000)
    assert rc == 0
    for (k, n, want, _), (st, r) in zip(rows, out):
        assert st == 0 and "%016X" % r.res64 == want, (k, n)
        assert r.fft_len == 49152 and r.maxerr < 0.4
    assert out[1][1].is_prp


def test_mixed_batch_all_256_vs_gmp_subset(golden):
    """The whole 256-test mixed batch (k odd in [3, 9999], n in [300K, 700K], FFT 20K-56K) in one
    run: every test completes its Gerbicz checks (d == x0 * t^(2^B) every M blocks, doubled tail)
    with round-off below 0.4, and the tests whose GMP RES64s are committed
    (tests/golden/res64_batch_subset.csv; GMP takes over 40 min per test at these sizes) match."""
    rows = golden("res64_batch_subset.csv")
    tests = workloads.batch_mixed()
    assert {(int(k), int(n)) for k, n, _, _ in rows} <= set(tests)  # a subset of the batch's tests
    say("mixed batch of", len(tests), "tests;", len(rows), "with GMP RES64s")
    rc, out = kite.batch_prp(tests, block=1000)
    assert rc == 0
    assert all(st == 0 and r.maxerr < 0.4 for st, r in out)
    assert len({r.fft_len for _, r in out}) >= 5  # mixed lengths advance together
    say("gerbicz errors", sum(r.gerbicz_errors for _, r in out), "retries", sum(r.roundoff_retries for _, r in out))
    res = {t: "%016X" % r.res64 for t, (st, r) in zip(tests, out)}
    for k, n, base, want in rows:
        assert res[(int(k), int(n))] == want, (k, n)


def test_batch_llr_matches_prst(golden):
    """LLR jobs in one batch (the squarings subtract 2 inside the column kernels): primes and the
    composite RES64s the reference checked against PRST 14.0.1888."""
    rows = golden("llr_expect.csv")
    jobs = [(kite.JOB_LLR, int(k), int(n), 0.0) for k, n, _ in rows]
    rc, out = kite.batch_run(jobs, block=1000)
    assert rc == 0
    for (k, n, want), r in zip(rows, out):
        say("llr", k, n, r.status, "%016X" % r.res64, r.confirm_runs, r.fft_len)
        assert r.status == 0 and r.maxerr < 0.4
        if want == "prime":
            assert r.is_prime and not r.factor
        else:
            assert not r.is_prime and "%016X" % r.res64 == want
            assert r.confirm_runs == 1  # a composite is confirmed once at the next length


@pytest.mark.parametrize("launch", [kite.LAUNCH_GRAPH, kite.LAUNCH_DIRECT])
def test_batch_mixed_fermat_and_llr(golden, launch):
    """Fermat and LLR jobs sharing launches, rounds replayed from graphs or launched directly
    (kite_run_opts.launch); each result equals the known one."""
    jobs = [(kite.JOB_PRP, 9357, 20000, 0.0), (kite.JOB_LLR, 9357, 20000, 0.0), (kite.JOB_PRP, 8191, 25000, 0.0),
            (kite.JOB_LLR, 3, 26459, 0.0), (kite.JOB_PRP, 3, 41628, 0.0), (kite.JOB_LLR, 33, 99999, 0.0)]
    rc, out = kite.batch_run(jobs, block=1000, launch=launch)
    assert rc == 0
    assert "%016X" % out[0].res64 == "79291CD2F47CE48C" and not out[0].is_prime
    assert "%016X" % out[1].res64 == "DC52061161A0DE1C" and not out[1].is_prime
    assert "%016X" % out[2].res64 == "D45538E6F07266C6"
    assert out[3].is_prime
    assert out[4].is_prime and out[4].res64 == 1
    assert "%016X" % out[5].res64 == "3FAF838186E8A836"


def test_batch_llr_fault_outvoted():
    """A corrupted first run of a prime's LLR job is outvoted by runs 1 and 2 (as kite_llr,
    tests/test_llr.c:69-74), while a neighbouring job is unaffected."""
    jobs = [(kite.JOB_LLR, 3, 41628, 0.0), (kite.JOB_LLR, 9357, 25000, 0.0)]
    rc, out = kite.batch_run(jobs, block=1000, fault_job=0, fault_at=20000)
    assert rc == 0
    assert out[0].is_prime and out[0].confirm_runs == 2
    assert "%016X" % out[1].res64 == "54A27E03D75F2447" and out[1].confirm_runs == 1


def test_serving_arrivals_slots_and_horizon():
    """Jobs start no earlier than they arrive, at most `slots` at a time, first come first served;
    a job the horizon cuts off reports KITE_JOB_PENDING."""
    jobs = [(kite.JOB_PRP, 9357, 20000, 0.0), (kite.JOB_LLR, 3, 26459, 0.3), (kite.JOB_PRP, 8191, 25000, 0.6),
            (kite.JOB_LLR, 9357, 20000, 0.9)]
    rc, out = kite.batch_run(jobs, slots=2, block=500)
    assert rc == 0
    for (kind, k, n, arr), r in zip(jobs, out):
        say(kind, k, n, "arrival", arr, "start %.3f finish %.3f" % (r.start, r.finish))
        assert r.status == 0 and r.start >= arr and r.finish > r.start
    assert "%016X" % out[0].res64 == "79291CD2F47CE48C"
    assert out[1].is_prime
    assert "%016X" % out[2].res64 == "D45538E6F07266C6"
    assert "%016X" % out[3].res64 == "DC52061161A0DE1C"
    # a long job and a short horizon: the short one finishes, the long one is pending
    rc, out = kite.batch_run([(kite.JOB_PRP, 9357, 20000, 0.0), (kite.JOB_PRP, 8045, 659321, 0.0)], horizon=2.0)
    assert out[0].status == 0 and "%016X" % out[0].res64 == "79291CD2F47CE48C"
    assert out[1].status == kite.JOB_PENDING and rc == kite.JOB_PENDING


def test_least_work_admission_order():
    """KITE_ORDER_LEAST_WORK with one slot: jobs that arrive together start in order of their work,
    squarings times FFT length (n L), and each result is the known one (tests/known_residues.csv)."""
    jobs = [(kite.JOB_PRP, 5, 30000, 0.0), (kite.JOB_PRP, 8191, 25000, 0.0), (kite.JOB_PRP, 9357, 20000, 0.0)]
    work = [n * kite.lib().kite_choose_fft_len(k, n) for _, k, n, _ in jobs]
    assert len(set(work)) == 3
    rc, out = kite.batch_run(jobs, slots=1, block=1000, order=kite.ORDER_LEAST_WORK)
    assert rc == 0
    order = sorted(range(3), key=lambda i: work[i])
    for a, b in zip(order, order[1:]):
        assert out[a].start <= out[a].finish <= out[b].start
    assert ["%016X" % r.res64 for r in out] == ["B7980452FDBC4183", "D45538E6F07266C6", "79291CD2F47CE48C"]


def test_hbm_budget_limits_jobs_in_flight():
    """With an HBM budget for two jobs, never more than two run at once; results unchanged."""
    L = kite.lib().kite_choose_fft_len(9357, 20000)
    budget = 2 * kite.lib().kite_job_bytes(kite.JOB_PRP, L) + 1
    jobs = [(kite.JOB_PRP, 9357, 20000, 0.0)] * 4
    rc, out = kite.batch_run(jobs, hbm_budget=budget, block=1000)
    assert rc == 0
    spans = sorted((r.start, r.finish) for r in out)
    for t, _ in spans:
        assert sum(1 for s, f in spans if s <= t < f) <= 2
    assert all("%016X" % r.res64 == "79291CD2F47CE48C" for r in out)


@pytest.mark.parametrize("parts,graphs,shape_parts", [(1, True, 0), (2, False, 0), (3, True, 0), (4, True, 0),
                                                     (1, False, 1), (2, True, 1), (3, True, 1)])
def test_parts_streams_ordering(parts, graphs, shape_parts):
    """Slots dealt to several streams (with shape parts: the most common shape's slots to `parts`
    streams and every other shape to a stream of its own while there are streams left):
    single-step and multi-step squarings, multiplies (whose launches read the slot table uploaded
    on the first stream) and a slot removed and re-added between calls; every value exact."""
    import ctypes
    cases = [(9357, 20000), (8191, 25000), (3, 26459), (7, 4001), (4523, 67231), (9357, 20011)]
    b = kite.Batch()
    r = random.Random(parts)
    try:
        assert b.parts(parts) == parts
        assert b.shape_parts(shape_parts) == shape_parts
        b.graphs(graphs)
        vals = {}
        for k, n in cases:
            s = b.add(k, n, nnums=2)
            x, y = r.randrange(k * (1 << n) - 1), r.randrange(k * (1 << n) - 1)
            b.set(s, 0, x)
            b.set(s, 1, y)
            vals[s] = [k, n, x, y]
        cap = b.lib.kite_batch_capacity(b.p)
        dst = (ctypes.c_int32 * cap)(*[0] * cap)
        src = (ctypes.c_int32 * cap)(*[1] * cap)
        for rnd in range(3):
            b.square(1)
            b.square(7)
            assert b.lib.kite_batch_mul(b.p, dst, src) == 0
            for s, v in vals.items():
                k, n = v[0], v[1]
                v[2] = mod_riesel(sqr_n(v[2], 8, k, n) * v[3], k, n)
            if rnd == 1:  # remove one slot and add a new test in its place
                s0 = next(iter(vals))
                b.lib.kite_batch_remove(b.p, s0)
                del vals[s0]
                k, n = 1203, 90001
                s = b.add(k, n, nnums=2)
                x, y = r.randrange(k * (1 << n) - 1), r.randrange(k * (1 << n) - 1)
                b.set(s, 0, x)
                b.set(s, 1, y)
                vals[s] = [k, n, x, y]
        for s, (k, n, x, y) in vals.items():
            assert b.get(s, 0) == x, (parts, k, n)
        assert max(b.maxerr().values()) < 0.4
    finally:
        b.close()
