"""Persistent launches (kite_batch_set_persistent): the square-mode row and fused column passes run
on about the device's resident capacity, each work-group walking a fixed list of jobs that the host
wrote once. Every value must be exact and bit-identical to the one-work-group-per-job launches,
whatever the streams, the XCD layout, idle slots and repeated graph replays."""
import ctypes
import random

import pytest

from ahnum_amd import kite, workloads
from test_gpu_engine import sqr_n

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_batch(tests, seed, parts, xcds, persistent):
    b = kite.Batch()
    b.parts(parts)
    b.xcds(xcds)
    assert b.persistent(persistent) == int(bool(persistent))
    r = random.Random(seed)
    vals = {}
    for k, n in tests:
        s = b.add(k, n)
        a = r.randrange(k * (1 << n) - 1)
        b.set(s, 0, a)
        vals[s] = (k, n, a)
    return b, vals


# tests at n~660K (L 49152, and test 31 of the seeded 40 at L 51200: the two-length fused column
# kernel and the register row kernel, both with persistent variants), plus small tests whose
# passes have none. Every value is compared with the direct launches; every fourth (and each small
# one) with exact squarings too.
LARGE = workloads.single_large(20) + [workloads.single_large(40)[31]]
MIX = LARGE + [(9357, 20000), (4523, 67231), (3, 26459)]


@pytest.mark.parametrize("parts,xcds", [(1, 8), (2, 8), (2, 1), (3, 5)])
def test_persistent_exact_and_same_as_direct(parts, xcds):
    say("parts", parts, "xcds", xcds)
    bp, vals = make_batch(MIX, 5, parts, xcds, True)
    bd, _ = make_batch(MIX, 5, parts, xcds, False)
    try:
        say("lengths", sorted({bp.fft_len(s) for s in vals}))
        for b in (bp, bd):
            b.square(5)  # a captured graph of 5 steps
            b.square(1)  # direct launches
            b.square(5)  # the graph replayed
        for s, (k, n, a) in vals.items():
            got = bp.get(s, 0)
            assert got == bd.get(s, 0), (k, n)
            if s % 4 == 0 or n < 100000:
                assert got == sqr_n(a, 11, k, n), (k, n, bp.fft_len(s))
        assert max(bp.maxerr().values()) < 0.4
    finally:
        bp.close()
        bd.close()


def test_persistent_idle_slots_and_counts():
    """Slots that sit out (which = -1) and slots with fewer steps than the call: their jobs are
    skipped inside the persistent loop, the others' values stay exact."""
    b, vals = make_batch(LARGE, 7, 2, 8, True)
    try:
        cap = b.lib.kite_batch_capacity(b.p)
        which = (ctypes.c_int32 * cap)(*[(-1 if s % 5 == 0 else 0) for s in range(cap)])
        counts = [1 + (s % 4) for s in range(cap)]
        cnt = (ctypes.c_uint32 * cap)(*counts)
        for _ in range(2):
            assert b.lib.kite_batch_square(b.p, which, cnt) == 0
        for s, (k, n, a) in vals.items():
            if s % 5 == 0:
                assert b.get(s, 0) == a, (s, k, n)
            elif s % 3 == 0:
                assert b.get(s, 0) == sqr_n(a, 2 * counts[s], k, n), (s, k, n)
    finally:
        b.close()


def test_persistent_timing_names():
    """Timed persistent launches are reported under their own kernel names."""
    b, vals = make_batch(LARGE, 9, 1, 8, True)
    try:
        b.timing(1)
        b.square(2)
        b.sync()
        names = [t[0] for t in b.kernel_times()]
        say(names)
        assert any(nm.startswith("k_cols_fused_rr_p<") for nm in names)
        assert "k_rows_sq256_p" in names
        b.timing(0)
        for s, (k, n, a) in list(vals.items())[:4]:
            assert b.get(s, 0) == sqr_n(a, 2, k, n)
    finally:
        b.close()
