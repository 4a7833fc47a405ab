"""ctypes binding to the CPU oracle (oracle/bin/liboracle.so) for tests and the benchmark's CPU
baseline only. The oracle restates the reference's portable algorithm; it is never the product."""
import ctypes
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# KITE_ORACLE_LIB: another build of the oracle (the benchmark's CPU baseline times the -O3 build)
PATH = os.environ.get("KITE_ORACLE_LIB") or os.path.join(ROOT, "oracle", "bin", "liboracle.so")

P = ctypes.c_void_p
u64, u32, i64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64
DP = ctypes.POINTER(ctypes.c_double)
LP = ctypes.POINTER(u64)


class PrpRes(ctypes.Structure):
    _fields_ = [("res64", u64), ("is_prp", ctypes.c_int), ("fft_len", u32), ("maxerr", ctypes.c_double),
                ("gerbicz_errors", u32), ("roundoff_retries", u32)]


class LlrRes(ctypes.Structure):
    _fields_ = [("is_prime", ctypes.c_int), ("factor", u64), ("res64", u64), ("P", u32), ("fft_len", u32),
                ("maxerr", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(PATH)
        sig = {
            "or_fft_len_ok": (ctypes.c_int, [u32]),
            "or_safe_bpw": (ctypes.c_double, [u32, u64]),
            "or_choose_fft_len": (u32, [u64, u64]),
            "or_next_fft_len": (u32, [u32]),
            "or_ctx_new": (P, [u64, u64, u32, ctypes.c_int]),
            "or_ctx_free": (None, [P]),
            "or_fft_len": (u32, [P]),
            "or_nlimbs": (ctypes.c_size_t, [P]),
            "or_layout": (None, [P, ctypes.POINTER(ctypes.c_uint8), DP, LP]),
            "or_s_at": (u64, [P, u64]),
            "or_weight": (ctypes.c_double, [P, u64, u64]),
            "or_set_limbs": (None, [P, DP, LP]),
            "or_get_limbs": (None, [P, DP, LP]),
            "or_set_u64": (None, [P, DP, u64]),
            "or_square": (None, [P, DP]),
            "or_square_n": (None, [P, DP, u64]),
            "or_mul": (None, [P, DP, DP]),
            "or_mul_small": (None, [P, DP, i64]),
            "or_add": (None, [P, DP, DP]),
            "or_sub": (None, [P, DP, DP]),
            "or_maxerr": (ctypes.c_double, [P]),
            "or_reset_maxerr": (None, [P]),
            "or_to_digits": (None, [P, DP, ctypes.POINTER(i64)]),
            "or_carry_digits": (None, [P, ctypes.POINTER(i64)]),
            "or_digits_to_limbs": (None, [P, ctypes.POINTER(i64), LP]),
            "or_residue_nm1": (ctypes.c_int, [LP, LP, ctypes.c_size_t, u64]),
            "or_prp": (ctypes.c_int, [u64, u64, u32, u32, u32, u64, ctypes.c_int, ctypes.POINTER(PrpRes)]),
            "or_llr_choose_P": (ctypes.c_int, [u64, u64, ctypes.POINTER(u32), LP]),
            "or_llr": (ctypes.c_int, [u64, u64, u32, ctypes.c_int, ctypes.POINTER(LlrRes)]),
            "or_c2p_shape": (ctypes.c_int, [u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
            "or_dft": (None, [DP, DP, DP, DP, u32, ctypes.c_int]),
            "or_sincos2pi": (None, [u64, u64, DP, DP]),
        }
        for name, (r, a) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def to_limbs(v, nl):
    v &= (1 << (64 * nl)) - 1
    return (u64 * nl).from_buffer_copy(v.to_bytes(8 * nl, "little"))


def from_limbs(a):
    return int.from_bytes(bytes(a), "little")


class OCtx:
    REF, C2P = 0, 1

    def __init__(self, k, n, L=0, backend=0):
        self.k, self.n, self.N = k, n, k * (1 << n) - 1
        self.p = lib().or_ctx_new(k, n, L, backend)
        if not self.p:
            raise RuntimeError("oracle ctx failed")
        self.L = lib().or_fft_len(self.p)
        self.nl = lib().or_nlimbs(self.p)

    def close(self):
        lib().or_ctx_free(self.p)

    def num(self, v=None):
        x = (ctypes.c_double * self.L)()
        if v is not None:
            self.set(x, v)
        return x

    def set(self, x, v):
        lib().or_set_limbs(self.p, x, to_limbs(v % self.N, self.nl))

    def get(self, x):
        a = (u64 * self.nl)()
        lib().or_get_limbs(self.p, x, a)
        return from_limbs(a)
