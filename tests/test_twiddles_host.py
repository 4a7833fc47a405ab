"""The library's own twiddle generator (no GPU): every device table (w_R, w_C, the H-sized split and
four-step tables, the two-level w_H table) is built from kite_debug_root_of_unity's values, so they
are checked here against mpmath at 40 digits, at the transform lengths the engine uses. Each value
must be within 0.51 ulp of the true cosine or sine (correctly rounded but for a double rounding
from long double), and the two-level product thi[t >> s] * tlo[t mod 2^s] the long shapes use
within 4e-16 of w_H^t."""
import ctypes
import math
import random

import pytest

mpmath = pytest.importorskip("mpmath")

from ahnum_amd import kite

mpmath.mp.dps = 40

# H = L/2 of benchmark and test lengths (49152, 51200, 204800, 442368, 1638400, 2359296), column
# and row lengths, and a power of two
LENGTHS = [24576, 25600, 102400, 221184, 819200, 1179648, 96, 100, 400, 256, 2048, 1 << 20]


def lib_root(j, n):
    L = kite.lib()
    f = L.kite_debug_root_of_unity
    f.restype = None
    f.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    c, s = ctypes.c_double(), ctypes.c_double()
    f(j, n, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def ulps(x, exact):
    """|x - exact| in units of the last place of the double nearest to exact."""
    e = float(exact)
    u = math.ulp(e) if e != 0 else math.ulp(0.0)
    if abs(e) > 0 and math.frexp(abs(e))[0] == 0.5 and abs(mpmath.mpf(x)) < abs(exact):
        u = u / 2  # just below a power of two the spacing halves
    return float(abs(mpmath.mpf(x) - exact) / u)


@pytest.mark.parametrize("n", LENGTHS)
def test_roots_within_half_ulp(n):
    r = random.Random(n)
    js = {0, 1, n // 8, n // 8 + 1, n // 4 - 1, n // 4, n // 2, 3 * n // 8, n - 1} | {r.randrange(n) for _ in range(60)}
    worst = 0.0
    for j in sorted(js):
        c, s = lib_root(j, n)
        if 4 * j % n == 0:  # quarter turns: exactly 0 and +-1 (mpmath's pi leaves ~1e-43 there)
            q = 4 * j // n
            assert (c, s) == ((1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0))[q], (n, j, c, s)
            continue
        x = 2 * mpmath.pi * j / n
        worst = max(worst, ulps(c, mpmath.cos(x)), ulps(s, mpmath.sin(x)))
    assert worst <= 0.51, (n, worst)


@pytest.mark.parametrize("H", [24576, 102400, 221184, 819200, 1179648])
def test_two_level_product(H):
    """w_H^t = thi[t >> sh] tlo[t mod 2^sh] (sh as the engine's shape tables choose it)."""
    sh = 0
    while (1 << (2 * sh)) < H:
        sh += 1
    r = random.Random(H)
    worst = 0.0
    for t in [0, 1, H - 1, (1 << sh) - 1, 1 << sh] + [r.randrange(H) for _ in range(200)]:
        hc, hs = lib_root((t >> sh) << sh, H)  # thi entry: w_H^(hi 2^sh)
        lc, ls = lib_root(t & ((1 << sh) - 1), H)  # tlo entry
        # the tables hold e^{-2 pi i t/H} = (c, -s); their product, as the kernels form it
        pr, pi = hc * lc - (-hs) * (-ls), hc * (-ls) + (-hs) * lc
        x = 2 * mpmath.pi * t / H
        worst = max(worst, float(abs(mpmath.mpf(pr) - mpmath.cos(x))), float(abs(mpmath.mpf(pi) + mpmath.sin(x))))
    assert worst <= 4e-16, (H, worst)
