"""Parity of the HIP engine (through the C-ABI library) with exact arithmetic, GMP vectors and the
reference's own residues. Values are compared exactly (bit-exact limbs / RES64)."""
import ctypes
import os
import random
import sys

import pytest

from ahnum_amd import kite, workloads
import oracle as O
import bigint

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mod_riesel(x, k, n):
    """x mod N for N = k*2^n - 1 by k*2^n == 1: x = h*2^n + l, h = q*k + r -> l + q + r*2^n."""
    N = k * (1 << n) - 1
    mask = (1 << n) - 1
    while x.bit_length() > n + k.bit_length() + 1:
        q, r = divmod(x >> n, k)
        x = (x & mask) + q + (r << n)
    while x >= N:
        x -= N
    while x < 0:
        x += N
    return x


def sqr_n(a, m, k, n):
    """a^(2^m) mod k*2^n-1 by m exact squarings."""
    for _ in range(m):
        a = mod_riesel(a * a, k, n)
    return a


def fnv64(v, nl):
    h = 0xCBF29CE484222325
    for byte in v.to_bytes(nl * 8, "little"):
        h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


ARITH_CASES = [
    (4523, 2000, 128), (9357, 5000, 384), (1, 3000, 256), (3, 2500, 192), (8915, 6000, 512),
    (7, 4001, 320), (9357, 20000, 0), (1, 44497, 0), (8191, 25000, 0), (4523, 67231, 0),
    (1203, 90001, 12800), (9999, 30011, 5120),  # ~6-7 bits/word: k spans several top digits
    (9001, 200000, 0), (8045, 659321, 0), (2009, 3322280, 0),
]


@pytest.mark.parametrize("k,n,L", ARITH_CASES)
def test_arith_vs_exact_ints(k, n, L):
    """The reference's test_gmp sequence (tests/test_gmp.c:99-160) on the GPU engine."""
    say("arith", k, n, L)
    r = random.Random(k * 7919 + n)
    with kite.Ctx(k, n, fft_len=L) as c:
        N = c.N
        a, b = r.randrange(N), r.randrange(N)
        x, y = c.num(a), c.num(b)
        assert x.get() == a
        x.square(20)
        t = sqr_n(a, 20, k, n)
        assert x.get() == t
        x.mul(y)
        t = mod_riesel(t * b, k, n)
        assert x.get() == t
        x.mul_small(3)
        t = t * 3 % N
        assert x.get() == t
        x.add(y)
        x.square()
        t = mod_riesel((t + b) ** 2, k, n)
        assert x.get() == t
        x.sub(y)
        t = (t - b) % N
        assert x.get() == t
        two = c.num(2)
        x.set(4)
        t = 4
        for _ in range(30 if n < 500000 else 3):
            x.square()
            x.sub(two)
            t = mod_riesel(t * t - 2, k, n)
        assert x.get() == t
        x.set(3)
        x.square(30)
        assert x.get() == sqr_n(3, 30, k, n)
        assert c.maxerr() < 0.4
        for v in (x, y, two):
            v.free()


@pytest.mark.parametrize("k,n,m", [
    (883, 4046207, 4), (109, 5665875, 3), (497, 5902972, 3),  # the 3M-6M configuration (L 294912-409600)
    (91, 24767982, 2),  # record size, L 1638400
])
def test_large_lengths_vs_exact_ints(k, n, m):
    """Squarings at the large FFT lengths the selector picks for n = 3M..25M, checked against exact
    integer squaring mod N."""
    L = kite.lib().kite_choose_fft_len(k, n)
    say("large", k, n, "L", L, "squarings", m)
    r = random.Random(k + n)
    with kite.Ctx(k, n) as c:
        assert c.L == L
        a = r.randrange(c.N)
        x = c.num(a)
        x.square(m)
        assert x.get() == sqr_n(a, m, k, n)
        assert c.maxerr() < 0.4
        x.free()


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_shape_variants_agree(order):
    """Every shape variant (R x C split) gives the same value (the autotuner may pick any)."""
    k, n = 9357, 369717
    say("variant", order)
    try:
        c = kite.Ctx(k, n, plan_order=order + 1)
    except RuntimeError:
        pytest.skip("no such variant")
    with c:
        a = random.Random(5).randrange(c.N)
        x = c.num(a)
        x.square(8)
        assert x.get() == sqr_n(a, 8, k, n)


def test_gpu_matches_oracle_digits_value():
    """Same input through the GPU and the oracle's two-pass backend: identical values after a
    mixed op sequence (the oracle is the checker)."""
    k, n = 9357, 30011
    r = random.Random(1)
    oc = O.OCtx(k, n, 0, O.OCtx.C2P)
    with kite.Ctx(k, n) as c:
        a, b = r.randrange(c.N), r.randrange(c.N)
        x, y = c.num(a), c.num(b)
        ox, oy = oc.num(a), oc.num(b)
        for step in range(6):
            x.square(7)
            O.lib().or_square_n(oc.p, ox, 7)
            x.mul(y)
            O.lib().or_mul(oc.p, ox, oy)
            assert x.get() == oc.get(ox)
    oc.close()


def test_chain_vectors_gmp(golden):
    """x_i = 3^(k 2^i) mod N against GMP (tests/golden/chain_x.csv), up to the record size
    91*2^24767982-1 (L = 1638400) at i = 50, 2000 and 10000."""
    rows = golden("chain_x.csv")
    ctxs = {}
    try:
        for k, n, base, i, low, fnv in rows:
            k, n, i = int(k), int(n), int(i)
            say("chain", k, n, i)
            if (k, n) not in ctxs:
                c = kite.Ctx(k, n)
                ctxs[(k, n)] = [c, c.num(pow(3, k, c.N)), 0]
            c, x, at = ctxs[(k, n)]
            x.square(i - at)
            ctxs[(k, n)][2] = i
            v = x.get()
            assert "%016X" % (v & 0xFFFFFFFFFFFFFFFF) == low
            assert "%016X" % fnv64(v, (n + 127) // 64) == fnv
    finally:
        for c, x, _ in ctxs.values():
            x.free()
            c.close()


def test_full_size_squarings_agree_across_lengths():
    """10^4 squarings of a seeded random residue mod 8045*2^659321-1 at L = 49152 and at the next
    length: the same value (an FFT fault at full size would not reproduce at a different length),
    and max round-off below 0.4."""
    k, n = workloads.single_large()[0]
    vals, errs = [], []
    for L in (0, 1):
        with kite.Ctx(k, n) as c0:
            Ls = c0.L
        if L:
            Ls = kite.lib().kite_next_fft_len(Ls)
        say("10^4 squarings at L", Ls)
        with kite.Ctx(k, n, fft_len=Ls) as c:
            a = workloads.random_residue_limbs(c.N, 2)
            x = c.num(a)
            c.reset_maxerr()
            x.square(10000)
            errs.append(c.maxerr())
            vals.append(x.get())
            x.free()
    assert vals[0] == vals[1]
    assert max(errs) < 0.4


def prp_row(k, n, want, **kw):
    say("prp", k, n, kw)
    rc, r = kite.prp(k, n, **kw)
    assert rc == 0
    assert "%016X" % r.res64 == want, (k, n)
    assert r.maxerr < 0.4
    return r


def test_prp_known_residues(golden):
    """The reference's run_known.sh rows with n <= 100000 (tests/known_residues.csv:2-13)."""
    for k, n, want, src in golden("known_residues.csv"):
        if int(n) <= 100000:
            r = prp_row(int(k), int(n), want)
            assert r.is_prp == (want == "0000000000000001")


def test_prp_seeded_small_vs_gmp(golden):
    """16 seeded Fermat tests, k odd in [1001, 9999], n in [60000, 70000], RES64 by GMP."""
    rows = golden("res64_plain_small.csv")
    assert [(int(k), int(n)) for k, n, _, _ in rows] == workloads.plain_small()
    for k, n, base, want in rows:
        prp_row(int(k), int(n), want, block=1000)


def test_gerbicz_catches_fault():
    """kite prp 1203 90001 --fault 50000 --expect 4BE9F1953E350085 (CMakeLists.txt:95)."""
    r = prp_row(1203, 90001, "4BE9F1953E350085", fault_at=50000)
    assert r.gerbicz_errors >= 1


def test_prp_reference_vector_200001():
    """9357*2^200001-1 -> BDD48C45E17CDD83 (the reference's tests/prst_kill.py:22-24)."""
    prp_row(9357, 200001, "BDD48C45E17CDD83")


def test_checkpoint_stop_and_resume(tmp_path):
    """Stop part way (abort flag from the progress callback), resume from the KITECKP1 file,
    same RES64 (the reference's prst_resume behaviour)."""
    ck = str(tmp_path / "t.ckpt")
    flag = ctypes.c_int(0)
    k, n = 9357, 20000

    def prog(i, nn, e):
        if i >= 8000:
            flag.value = 1

    rc, r = kite.prp(k, n, checkpoint=ck, abort_flag=flag, progress=prog, block=100)
    assert rc == -5 and os.path.exists(ck)
    it = ctypes.c_uint64()
    assert kite.lib().kite_prp_checkpoint_iter(ck.encode(), k, n, 3, ctypes.byref(it)) == 0
    assert it.value >= 8000
    flag.value = 0
    rc, r = kite.prp(k, n, checkpoint=ck, block=100)
    assert rc == 0 and r.resumed and r.resumed_at >= 8000
    assert "%016X" % r.res64 == "79291CD2F47CE48C"


def test_llr_expectations(golden):
    """LLR primes and composite RES64s checked by the reference against PRST 14.0.1888."""
    for k, n, want in golden("llr_expect.csv"):
        say("llr", k, n)
        rc, r = kite.llr(int(k), int(n))
        assert rc == 0
        assert r.maxerr < 0.4
        if want == "prime":
            assert r.is_prime and not r.factor
        else:
            assert not r.is_prime and "%016X" % r.res64 == want


def test_llr_fault_outvoted():
    """A prime whose run 0 is corrupted is settled by runs 1 and 2 (tests/test_llr.c:69-74)."""
    rc, r = kite.llr(3, 41628, inject_fault_at=20000)
    assert rc == 0 and r.is_prime and r.confirm_runs == 2


def test_llr_impossible_roundoff_limit():
    """A round-off limit nothing meets gives up with -2 (tests/test_llr.c:135-139)."""
    rc, r = kite.llr(3, 41628, max_roundoff=1e-9)
    assert rc == -2


# column lengths with odd parts 9, 15, 25, 27 (compile-time kernels): the length whose preferred shape
# has that column length (C = 4R or so), squared through the batch's fused column kernel
ODD_PART_LENGTHS = [(36, 9216), (60, 30720), (72, 36864), (100, 51200), (108, 55296), (120, 122880),
                    (144, 147456), (200, 409600), (216, 442368), (240, 491520), (288, 589824),
                    (400, 1638400), (432, 1769472), (480, 1966080), (576, 2359296),
                    # and the other lengths R 288..512 of 8-column groups (C = 2048)
                    (320, 1310720), (384, 1572864), (448, 1835008), (512, 2097152)]


@pytest.mark.parametrize("R,L", ODD_PART_LENGTHS)
def test_odd_part_column_lengths_exact(R, L):
    """Two squarings at each new column length against exact integers, through its own
    compile-time kernels (named in the launch timings)."""
    k = 3
    n = int(L * 13.5)
    say("odd part", R, "L", L, "n", n)
    r = random.Random(R)
    b = kite.Batch()
    try:
        s = b.add(k, n, fft_len=L)
        a = r.randrange(k * (1 << n) - 1)
        b.set(s, 0, a)
        b.timing(1)
        b.square(2)
        b.sync()
        names = [name for name, *_ in b.kernel_times()]
        say("  kernels", names)
        assert any(name.endswith("<%d>" % R) for name in names), names
        assert b.get(s, 0) == bigint.sqr_n(a, 2, k, n)
        assert b.maxerr()[s] < 0.4
    finally:
        b.close()


def test_adjacent_odd_part_lengths_share_launches():
    """Tests at L = 49152 (R 96) and 51200 (R 100) in one batch (one stream, so that all three meet
    in the same launches) run in the pair kernels <96, 100>."""
    r = random.Random(96)
    b = kite.Batch()
    try:
        b.parts(1)
        vals = []
        for k, n, L in [(8045, 659321, 49152), (9999, 670000, 51200), (3, 660001, 49152)]:
            s = b.add(k, n, fft_len=L)
            a = r.randrange(k * (1 << n) - 1)
            b.set(s, 0, a)
            vals.append((s, k, n, a))
        b.timing(1)
        b.square(2)
        b.sync()
        names = [name for name, *_ in b.kernel_times()]
        say("kernels", names)
        assert any("_rr<96,100>" in name for name in names), names
        for s, k, n, a in vals:
            assert b.get(s, 0) == bigint.sqr_n(a, 2, k, n), (k, n)
    finally:
        b.close()
