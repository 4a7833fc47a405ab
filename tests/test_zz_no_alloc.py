"""kite_square_n_static, the squaring entry point that allocates nothing (its device scratch is a
static array of the library), against kite_square_n on the same inputs: bit-exact limbs and the
same round-off, on two calls in a row, the second with a smaller number (the scratch keeps the
first call's contents), and a length whose carry flags do not fit the scratch is refused."""
import ctypes
import random
import sys

import pytest

from ahnum_amd import kite

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lib():
    L = kite.lib()
    L.kite_square_n_static.restype = ctypes.c_int
    L.kite_square_n_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    return L


def square_both(k, n, count, seed, fft_len=0):
    """count squarings of one random residue mod k*2^n-1 by both entry points; their limbs and
    round-off, and the entry point's return code."""
    r = random.Random(seed)
    with kite.Ctx(k, n, fft_len=fft_len) as c:
        a = r.randrange(c.N)
        x, y = c.num(a), c.num(a)
        c.reset_maxerr()
        x.square(count)  # kite_square_n
        want, err_want = x.get(), c.maxerr()
        c.reset_maxerr()
        rc = lib().kite_square_n_static(c.ptr, y.p, count)
        got, err_got = (y.get(), c.maxerr()) if rc == 0 else (None, None)
        x.free()
        y.free()
        return c.L, rc, want, got, err_want, err_got


@pytest.mark.parametrize("k,n,count", [(8045, 659321, 100), (9357, 20000, 200)])
def test_static_squarings_match(k, n, count):
    say("static squarings", k, n, count)
    L, rc, want, got, ew, eg = square_both(k, n, count, seed=k + n)
    say("L", L, "rc", rc, "maxerr", ew, eg)
    assert rc == 0
    assert got == want
    assert eg == ew and eg < 0.4


def test_two_calls_second_smaller():
    """A long-length call, then a short one: the scratch still holds the first call's flags."""
    L1, rc1, w1, g1, _, _ = square_both(8045, 659321, 64, seed=1)
    L2, rc2, w2, g2, _, _ = square_both(9357, 30011, 64, seed=2)
    say("lengths", L1, L2)
    assert L2 < L1
    assert rc1 == 0 and g1 == w1
    assert rc2 == 0 and g2 == w2


def test_too_long_for_the_scratch_is_refused():
    """Length 3145728 (768 x 2048 complex, 4 columns per group): 393216 carry flags, more than the
    262112 the scratch holds: refused with -2 before anything runs, and the context still works."""
    k, n, L = 3, 40000001, 3145728
    say("refusal", k, n, L)
    with kite.Ctx(k, n, fft_len=L) as c:
        assert c.L == L
        x = c.num(3)
        rc = lib().kite_square_n_static(c.ptr, x.p, 2)
        assert rc == -2
        x.square(2)
        assert x.get() == 3 ** 4
        x.free()
