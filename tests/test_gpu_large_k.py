"""Large k on the GPU: the carries between work-groups and the top fold at the reference's k limit
(k < 2^30, src/ctx.c:225), against exact integers.

With k near 2^29 the products before the carry reach about 2^50 at the selected lengths, so a row
chunk's carry reaches 2^44 (2^39 for k = 2^25 + 1 at b0 = 8): the fused column pass must carry it
through its 48-bit flag field, and the split passes (digits below 6 bits) through their int64 carries.
A forced length whose products would pass 2^53 is refused at kite_ctx_new.
"""
import math
import random
import sys

import pytest

from ahnum_amd import kite
import bigint

pytestmark = pytest.mark.gpu

K29 = (1 << 29) - 1
K25 = (1 << 25) + 1
K30 = (1 << 30) - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# (k, n, forced length or 0): default lengths with fused (b0 >= 6) and split (b0 < 6) column
# passes, and forced longer lengths down to 2-4 bits per digit (the top region then spans 8-13 digits)
CASES = [
    (K29, 100003, 0),       # L 16384, b0 6: fused pass, carries ~2^44
    (K25, 100003, 0),       # L 12288, b0 8: fused pass
    (K30, 20011, 0),        # L 3584, b0 5: split passes
    (999999937, 1000, 0),   # L 128: the largest products the selector allows (~2^51.4)
    (K29, 659321, 0),       # L 131072, b0 5: split passes
    (K29, 3322280, 0),      # L 786432, b0 4: needs 8-element carry sub-chunks (R 192, not 384)
    (K29, 100003, 28672),   # forced: b0 3
    (K25, 100003, 20480),   # forced: b0 4
    (K25, 100003, 40960),   # forced: b0 2, the leftover of each absorbed carry just below 2^(b0-1)
    (K29, 100003, 36864),   # forced: b0 2 (2.71 bits per digit), generic column kernels
]


@pytest.mark.parametrize("k,n,L", CASES)
def test_large_k_squarings_exact(k, n, L):
    say("large k", k, n, "forced L" if L else "default L", L)
    r = random.Random(k ^ n)
    with kite.Ctx(k, n, fft_len=L) as c:
        say("  L", c.L, "bits/word %.3f" % ((n + math.log2(k)) / c.L))
        a, b = r.randrange(c.N), r.randrange(c.N)
        x, y = c.num(a), c.num(b)
        assert x.get() == a
        x.square(6)
        t = bigint.sqr_n(a, 6, k, n)
        assert x.get() == t
        x.mul(y)
        t = bigint.mod_riesel(bigint.mul(t, b), k, n)
        assert x.get() == t
        x.square(1)
        assert x.get() == bigint.sqr_n(t, 1, k, n)
        # balanced extremes: N - 1 = -1 and (N - 1) / 2 squared
        x.set(c.N - 1)
        x.square(2)
        assert x.get() == 1
        h = (c.N - 1) // 2
        x.set(h)
        x.square(1)
        assert x.get() == bigint.sqr_n(h, 1, k, n)
        say("  maxerr %.4f" % c.maxerr())
        assert c.maxerr() < 0.4
        for v in (x, y):
            v.free()


def test_large_k_batch_exact():
    """Tests at k = 2^29 - 1, 2^25 + 1, 2^30 - 1 and small k in one batch: two squarings of each slot
    through the shared launches, against exact squaring."""
    tests = [(K29, 100003), (K25, 100003), (3, 100003), (8045, 100003), (K30, 100003), (K29, 99989)]
    r = random.Random(29)
    b = kite.Batch()
    try:
        vals = []
        for k, n in tests:
            s = b.add(k, n)
            a = r.randrange(k * (1 << n) - 1)
            b.set(s, 0, a)
            vals.append((s, k, n, a))
        say("batch lengths", [(k, n, b.fft_len(s)) for s, k, n, _ in vals])
        b.square(2)
        b.sync()
        errs = b.maxerr()
        for s, k, n, a in vals:
            assert b.get(s, 0) == bigint.sqr_n(a, 2, k, n), (k, n)
            assert errs[s] < 0.4, (k, n)
    finally:
        b.close()


@pytest.mark.parametrize("k,n,L", [
    (K29, 100003, 12288),   # 8.1 bits per digit: products would reach 2^53.8
    (K29, 100003, 40960),   # 2.44 bits per digit: carries of 2^40.7 outgrow 15 digits (diverged, round-off 0.5)
    (K25, 100003, 49152),   # 2.04 bits per digit: the same
])
def test_unsafe_forced_lengths_refused(k, n, L):
    """Forced lengths the engine cannot run exactly are refused at kite_ctx_new; the selector's own
    length for the same number is accepted."""
    with pytest.raises(RuntimeError):
        kite.Ctx(k, n, fft_len=L)
    with kite.Ctx(k, n) as c:
        assert c.L == kite.lib().kite_choose_fft_len(k, n)


@pytest.mark.parametrize("k,n", [(K29, 5003), (K25, 20011), (K30, 4001)])
def test_large_k_prp_and_llr_exact(k, n):
    """kite_prp's RES64 is the low word of 3^(N-1) mod N; kite_llr's is that of the Lucas V term
    (zero iff prime), both from exact integers."""
    N = k * (1 << n) - 1
    rc, r = kite.prp(k, n)
    y = pow(3, N - 1, N)
    say("prp", k, n, rc, "%016X" % r.res64, r.fft_len, "maxerr %.4f" % r.maxerr)
    assert rc == 0 and r.res64 == y & 0xFFFFFFFFFFFFFFFF and bool(r.is_prp) == (y == 1)
    rc, l = kite.llr(k, n)
    say("llr", k, n, rc, "%016X" % l.res64, "P", l.P, "factor", l.factor)
    assert rc == 0
    if l.factor:
        assert N % l.factor == 0
        return
    v = bigint.lucas_v_matrix(k, n, l.P)
    assert l.res64 == v & 0xFFFFFFFFFFFFFFFF and bool(l.is_prime) == (v == 0)
