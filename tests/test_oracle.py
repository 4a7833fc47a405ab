"""Pins the CPU oracle (oracle/kite_oracle.c) before it is trusted as the GPU engine's checker.

Each check compares against something the oracle does not compute itself: exact Python integer
arithmetic (pow, %), GMP golden vectors under tests/golden/, the reference's own residue
fixtures, closed forms of the layout (exact rationals), or a brute-force DFT. CPU only.
"""
import ctypes
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.filterwarnings("ignore")


def rng_cases(seed, count, nmax):
    r = random.Random(seed)
    out = []
    for t in range(count):
        k = r.randrange(1, 10000) | 1
        if t == 0:
            k = 1
        if t == 1:
            k = 9357
        n = 600 + r.randrange(nmax)
        L = 0
        if t % 2:  # force a low bits/word so k spans several top digits (tests/test_gmp.c:72-80)
            L = (n // (6 + r.randrange(6)) + 63) // 64 * 64
            while not O.lib().or_fft_len_ok(L):
                L += 64
        out.append((k, n, L))
    return out


@pytest.mark.parametrize("backend", [O.OCtx.REF, O.OCtx.C2P])
@pytest.mark.parametrize("k,n,L", rng_cases(11, 10, 3000) + rng_cases(12, 4, 9000))
def test_arith_vs_python_ints(backend, k, n, L):
    """The test_gmp.c sequence (tests/test_gmp.c:99-160) against exact integer arithmetic."""
    try:
        c = O.OCtx(k, n, L, backend)
    except RuntimeError:
        pytest.skip("no two-pass shape for this length")
    N = c.N
    r = random.Random(k * 1000003 + n)
    a, b = r.randrange(N), r.randrange(N)
    x, y = c.num(a), c.num(b)
    assert c.get(x) == a
    O.lib().or_square_n(c.p, x, 20)
    t = pow(a, 1 << 20, N)
    assert c.get(x) == t
    O.lib().or_mul(c.p, x, y)
    t = t * b % N
    assert c.get(x) == t
    O.lib().or_mul_small(c.p, x, 3)
    t = t * 3 % N
    assert c.get(x) == t
    O.lib().or_add(c.p, x, y)
    O.lib().or_square(c.p, x)
    t = (t + b) ** 2 % N
    assert c.get(x) == t
    O.lib().or_sub(c.p, x, y)
    t = (t - b) % N
    assert c.get(x) == t
    # LLR-style chain u <- u^2 - 2 from 4, 30 times
    two = c.num(2)
    O.lib().or_set_u64(c.p, x, 4)
    t = 4
    for _ in range(30):
        O.lib().or_square(c.p, x)
        O.lib().or_sub(c.p, x, two)
        t = (t * t - 2) % N
    assert c.get(x) == t
    assert O.lib().or_maxerr(c.p) < 0.4
    c.close()


@pytest.mark.parametrize("k,n", [(1, 1000), (3, 2001), (1365, 5000), (9357, 20000), (8191, 25000), (9999, 60000)])
def test_layout_closed_forms(k, n):
    """s_j = ceil(j T'/L) and w_j = 2^(s_j - j T'/L) against exact rationals; bits in {b, b+1};
    s_L = n + ceil(log2 k); lam = frac(log2 k) to 2^-60 (src/ctx.c:150-221)."""
    c = O.OCtx(k, n)
    L = c.L
    bits = (ctypes.c_uint8 * L)()
    w = (ctypes.c_double * L)()
    info = (ctypes.c_uint64 * 5)()
    O.lib().or_layout(c.p, bits, w, info)
    top_t, top_st, top_m, sL, lam = list(info)
    ik = k.bit_length() - 1
    frac = math.log2(k) - ik
    assert abs(lam / 2.0 ** 64 - frac) < 2e-15
    T = Fraction(n + ik) + Fraction(lam, 1 << 64)
    assert sL == n + (0 if k == 1 else math.ceil(math.log2(k)))
    b0 = math.floor(T / L)
    rs = random.Random(n)
    for j in [0, 1, 2, L - 1, L // 2] + [rs.randrange(L) for _ in range(40)]:
        sj = math.ceil(T * j / L)
        assert O.lib().or_s_at(c.p, j) == sj
        assert bits[j] in (b0, b0 + 1)
        e = sj - T * j / L
        assert 0 <= e < 1
        exact = 2.0 ** float(e)
        assert 1.0 <= w[j] < 2.0
        assert abs(w[j] - exact) <= 2 * math.ulp(exact)
    assert sum(bits) == sL
    assert top_t == L or top_st + top_m == n
    c.close()


def test_sincos_exact_reduction():
    import mpmath as mp
    mp.mp.dps = 40
    for n in [3, 5, 7, 8, 64, 96, 1000, 24576, 49152]:
        for j in list(range(0, min(n, 20))) + [n // 8, n // 4, n // 2, 3 * n // 4, n - 1]:
            cc, ss = ctypes.c_double(), ctypes.c_double()
            O.lib().or_sincos2pi(j, n, ctypes.byref(cc), ctypes.byref(ss))
            ec, es = mp.cos(2 * mp.pi * j / n), mp.sin(2 * mp.pi * j / n)
            # exact zeros of cos/sin come out as 0; mpmath leaves ~1e-40 there
            assert abs(cc.value - float(ec)) <= max(math.ulp(float(ec)) * 0.51, 1e-30)
            assert abs(ss.value - float(es)) <= max(math.ulp(float(es)) * 0.51, 1e-30)


def test_dft_brute_force_vs_numpy():
    r = np.random.default_rng(3)
    for n in [2, 3, 5, 7, 8, 12, 30]:
        re, im = r.standard_normal(n), r.standard_normal(n)
        ore, oim = np.zeros(n), np.zeros(n)
        dp = ctypes.POINTER(ctypes.c_double)
        O.lib().or_dft(re.ctypes.data_as(dp), im.ctypes.data_as(dp), ore.ctypes.data_as(dp),
                       oim.ctypes.data_as(dp), n, 0)
        ref = np.fft.fft(re + 1j * im)
        assert np.abs(ore + 1j * oim - ref).max() < 1e-12


def test_residue_nm1_and_crt():
    """base^(N+1) -> base^(N-1) mod N, including gcd(3, N) > 1 (src/prp.c:63-144)."""
    for k, n in [(9357, 200), (8191, 250), (5, 301), (3, 100), (2, 1) and (7, 401), (11, 333)]:
        N = k * (1 << n) - 1
        nl = (n + 127) // 64
        y = O.to_limbs(pow(3, N + 1, N), nl)
        assert O.lib().or_residue_nm1(y, O.to_limbs(N, nl), nl, 3) == 0
        assert O.from_limbs(y) == pow(3, N - 1, N)


@pytest.mark.parametrize("backend", [O.OCtx.REF, O.OCtx.C2P])
def test_prp_driver_small(backend):
    """Gerbicz driver end to end at small n, against pow(3, N-1, N), with and without a fault."""
    for k, n, fault in [(9357, 3000, 0), (5, 3001, 0), (8191, 2500, 1200), (1203, 4001, 2000)]:
        N = k * (1 << n) - 1
        r = O.PrpRes()
        rc = O.lib().or_prp(k, n, 3, 0, 40, fault, backend, ctypes.byref(r))
        assert rc == 0
        assert r.res64 == pow(3, N - 1, N) & 0xFFFFFFFFFFFFFFFF
        if fault:
            assert r.gerbicz_errors >= 1


def test_prp_known_residue_row(golden):
    """The reference's GMP row 9357*2^20000-1 (tests/known_residues.csv:2) through the oracle."""
    rows = {(int(k), int(n)): v for k, n, v, _ in golden("known_residues.csv")}
    r = O.PrpRes()
    assert O.lib().or_prp(9357, 20000, 3, 0, 0, 0, O.OCtx.C2P, ctypes.byref(r)) == 0
    assert "%016X" % r.res64 == rows[(9357, 20000)]
    assert r.maxerr < 0.4


def test_chain_vectors_gmp(golden):
    """x_i = 3^(k 2^i) mod N from GMP (tests/golden/chain_x.csv), oracle from x_0 = 3^k."""
    rows = [r for r in golden("chain_x.csv") if int(r[1]) < 100000][:10]
    done = {}
    for k, n, base, i, low, fnv in rows:
        k, n, i = int(k), int(n), int(i)
        key = (k, n)
        if key not in done:
            c = O.OCtx(k, n, 0, O.OCtx.C2P)
            x = c.num(pow(3, k, c.N))
            done[key] = (c, x, 0)
        c, x, at = done[key]
        O.lib().or_square_n(c.p, x, i - at)
        done[key] = (c, x, i)
        v = c.get(x)
        assert "%016X" % (v & 0xFFFFFFFFFFFFFFFF) == low
        nl = (n + 127) // 64
        h = 0xCBF29CE484222325
        for byte in v.to_bytes(nl * 8, "little"):
            h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
        assert "%016X" % h == fnv


def lucas_v(k, n, P):
    """V_{k 2^(n-1)}(P, -1) mod N by exact integers (the definition, src/llr.c:4-7)."""
    N = k * (1 << n) - 1
    v0, v1 = 2, P  # V_0, V_1 with Q = -1: V_{m+1} = P V_m + V_{m-1}
    for _ in range(k - 1):
        v0, v1 = v1, (P * v1 + v0) % N
    v = v1  # V_k
    v = (v * v + 2) % N  # k odd: V_2k = V_k^2 + 2
    for _ in range(n - 2):
        v = (v * v - 2) % N
    return v


def jacobi(a, m):
    a %= m
    r = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if m % 8 in (3, 5):
                r = -r
        a, m = m, a
        if a % 4 == 3 and m % 4 == 3:
            r = -r
        a %= m
    return r if m == 1 else 0


@pytest.mark.parametrize("k,n", [(3, 2000), (9357, 1500), (33, 2222), (127, 1201)])
def test_llr_oracle_vs_definition(k, n):
    N = k * (1 << n) - 1
    P, f = ctypes.c_uint32(), ctypes.c_uint64()
    assert O.lib().or_llr_choose_P(k, n, ctypes.byref(P), ctypes.byref(f)) == 0
    p = 1
    while jacobi(p * p + 4, N) == 1:
        p += 1
    assert P.value == p
    if f.value:
        return
    r = O.LlrRes()
    assert O.lib().or_llr(k, n, 0, O.OCtx.C2P, ctypes.byref(r)) == 0
    v = lucas_v(k, n, p)
    assert r.res64 == v & 0xFFFFFFFFFFFFFFFF
    assert r.is_prime == (v == 0)


def test_fft_length_rule_matches_reference_fixture(golden):
    """The reference's own length choice for 7 production cases (tests/large_fft_cases.csv:4-10)."""
    for k, n, _, L, _ in golden("large_fft_cases.csv"):
        assert O.lib().or_choose_fft_len(int(k), int(n)) == int(L)
    assert O.lib().or_choose_fft_len(9001, 662695) == 49152


@pytest.mark.parametrize("k,n,P", [(1011, 61, 5), (3, 200, 3), (9357, 130, 7), (127, 97, 1)])
def test_lucas_matrix_form_matches_recurrence(k, n, P):
    """The O(log k) matrix form the large-k GPU tests use, against the plain recurrence."""
    import bigint
    assert bigint.lucas_v_matrix(k, n, P) == lucas_v(k, n, P)
