"""Seeded synthetic workloads for the Fermat/LLR squaring engine.

Each workload is a list of (k, n) Riesel numbers N = k*2^n - 1, drawn from a xorshift64
generator (the generator of the reference's tests/test_gmp.c:20-26) so that the benchmark, the
golden-fixture script and the tests draw identical numbers from a seed.

  plain_small   16 Fermat base-3 tests, k odd in [1001, 9999], n in [60000, 70000]
  single_large  1 Fermat base-3 test, k odd in [1001, 9999], n in [650000, 670000]
  batch_mixed   256 Fermat base-3 tests, k odd in [3, 9999], n in [300000, 700000]
  large_fft     32 Fermat base-3 tests, k odd in [3, 999], n in [3M, 6M]
  record        1 test, k odd < 100, n in [24M, 25M]
  serve_stream  Poisson arrivals (mean 20/s) of Fermat and LLR tests, n log-uniform [200K, 5M]
"""
import math

MASK64 = (1 << 64) - 1


class XorShift64:
    """xorshift64 with shifts 13, 7, 17 (tests/test_gmp.c:20-26 of the reference)."""

    def __init__(self, seed):
        self.s = (seed or 0x9E3779B97F4A7C15) & MASK64

    def next(self):
        s = self.s
        s ^= (s << 13) & MASK64
        s ^= s >> 7
        s ^= (s << 17) & MASK64
        self.s = s
        return s

    def uniform_int(self, lo, hi):
        """Integer uniform in [lo, hi] (modulo bias below 2^-40 for these ranges)."""
        return lo + self.next() % (hi - lo + 1)

    def odd(self, lo, hi):
        """Odd integer uniform over the odd values in [lo, hi]."""
        lo |= 1
        if hi % 2 == 0:
            hi -= 1
        return lo + 2 * (self.next() % ((hi - lo) // 2 + 1))

    def unit(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def plain_small(count=16, seed=1):
    g = XorShift64(seed)
    return [(g.odd(1001, 9999), g.uniform_int(60000, 70000)) for _ in range(count)]


def single_large(count=1, seed=2):
    g = XorShift64(seed)
    return [(g.odd(1001, 9999), g.uniform_int(650000, 670000)) for _ in range(count)]


def batch_mixed(count=256, seed=3):
    g = XorShift64(seed)
    return [(g.odd(3, 9999), g.uniform_int(300000, 700000)) for _ in range(count)]


def large_fft(count=32, seed=4):
    g = XorShift64(seed)
    return [(g.odd(3, 999), g.uniform_int(3000000, 6000000)) for _ in range(count)]


def record(seed=5):
    g = XorShift64(seed)
    return [(g.odd(3, 99), g.uniform_int(24000000, 25000000))]


def serve_stream(count=64, rate=20.0, seed=6, n_lo=200000, n_hi=5000000, llr_frac=0.5):
    """Finite request stream: list of (arrival_s, kind, k, n) with Poisson arrivals."""
    g = XorShift64(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(1.0 - g.unit()) / rate
        kind = "llr" if g.unit() < llr_frac else "prp"
        k = g.odd(3, 9999)
        n = int(round(math.exp(math.log(n_lo) + g.unit() * (math.log(n_hi) - math.log(n_lo)))))
        out.append((t, kind, k, n))
    return out


def random_residue_limbs(N, seed):
    """A residue uniform mod N from xorshift64 limbs (as the reference's test_gmp.c draws them)."""
    g = XorShift64(seed)
    nl = (N.bit_length() + 63) // 64 + 1
    limbs = b"".join(g.next().to_bytes(8, "little") for _ in range(nl))  # limb i = i-th draw
    return int.from_bytes(limbs, "little") % N
