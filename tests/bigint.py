"""Exact big-integer arithmetic for the tests: products of multi-million-bit numbers through the
system's GMP (libgmp.so.10, loaded with ctypes), else Python's own integers. Test infrastructure:
the checker, never the thing measured."""
import ctypes
import ctypes.util


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def _load():
    for name in ("libgmp.so.10", ctypes.util.find_library("gmp")):
        if not name:
            continue
        try:
            g = ctypes.CDLL(name)
        except OSError:
            continue
        P = ctypes.POINTER(_Mpz)
        g.__gmpz_init.argtypes = [P]
        g.__gmpz_clear.argtypes = [P]
        g.__gmpz_import.argtypes = [P, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                    ctypes.c_size_t, ctypes.c_void_p]
        g.__gmpz_export.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_size_t,
                                    ctypes.c_int, ctypes.c_size_t, P]
        g.__gmpz_export.restype = ctypes.c_void_p
        g.__gmpz_mul.argtypes = [P, P, P]
        return g
    return None


_gmp = _load()


def _to_mpz(v, z):
    b = v.to_bytes((v.bit_length() + 7) // 8 or 1, "little")
    _gmp.__gmpz_import(ctypes.byref(z), len(b), -1, 1, 0, 0, b)


def _from_mpz(z):
    nbytes = (abs(z.size) * 64 + 7) // 8 + 8
    buf = ctypes.create_string_buffer(nbytes)
    count = ctypes.c_size_t(0)
    _gmp.__gmpz_export(buf, ctypes.byref(count), -1, 1, 0, 0, ctypes.byref(z))
    return int.from_bytes(buf.raw[:count.value], "little")


def mul(a, b):
    """a * b for non-negative integers."""
    if _gmp is None or a.bit_length() < 100000:
        return a * b
    x, y, r = _Mpz(), _Mpz(), _Mpz()
    for z in (x, y, r):
        _gmp.__gmpz_init(ctypes.byref(z))
    try:
        _to_mpz(a, x)
        _to_mpz(b, y)
        _gmp.__gmpz_mul(ctypes.byref(r), ctypes.byref(x), ctypes.byref(y))
        return _from_mpz(r)
    finally:
        for z in (x, y, r):
            _gmp.__gmpz_clear(ctypes.byref(z))


def mod_riesel(x, k, n):
    """x mod N for N = k*2^n - 1, by k*2^n == 1: x = h*2^n + l, h = q*k + r -> l + q + r*2^n."""
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
        a = mod_riesel(mul(a, a), k, n)
    return a


def lucas_v_matrix(k, n, P):
    """V_{k 2^(n-1)}(P, -1) mod N = k*2^n-1 from the definition: (V_{m+1}, V_m) = M^m (P, 2) with
    M = [[P, 1], [1, 0]] (V_{m+1} = P V_m + V_{m-1}), then V_2k = V_k^2 + 2 (k odd) and n - 2
    doublings V_2m = V_m^2 - 2. O(log k) products, so k up to 2^30 costs nothing."""
    N = k * (1 << n) - 1

    def mat_mul(A, B):
        return [[(A[0][0] * B[0][0] + A[0][1] * B[1][0]) % N, (A[0][0] * B[0][1] + A[0][1] * B[1][1]) % N],
                [(A[1][0] * B[0][0] + A[1][1] * B[1][0]) % N, (A[1][0] * B[0][1] + A[1][1] * B[1][1]) % N]]

    R, M, e = [[1, 0], [0, 1]], [[P, 1], [1, 0]], k
    while e:
        if e & 1:
            R = mat_mul(R, M)
        M = mat_mul(M, M)
        e >>= 1
    v = (R[1][0] * P + R[1][1] * 2) % N  # V_k: row 2 of M^k applied to (V_1, V_0)
    v = (v * v + 2) % N
    for _ in range(n - 2):
        v = (v * v - 2) % N
    return v
