"""ctypes binding to the MI355X kite library (ahnum_amd/lib/libkite.so).

Mirrors the reference C API (include/kite.h) one to one for tests and the benchmark. The library
is the product: there is no Python or CPU fallback, and loading fails loudly when it is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libkite.so")
BIN_PATH = os.path.join(HERE, "bin", "kite")

c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_p = ctypes.c_void_p

KITE_BACKEND_AUTO = 0
KITE_BACKEND_HIP = 5


class KiteOpts(ctypes.Structure):
    _fields_ = [("backend", ctypes.c_int), ("fft_len", c_u32), ("threads", c_u32), ("plan_order", c_u32)]


PROGRESS_FN = ctypes.CFUNCTYPE(None, c_p, c_u64, c_u64, c_dbl)
POINT_FN = ctypes.CFUNCTYPE(ctypes.c_int, c_p, c_u32, c_u64, ctypes.POINTER(c_u64))
CONFIRM_FN = ctypes.CFUNCTYPE(None, c_p, ctypes.c_int)


class PrpOpts(ctypes.Structure):
    _fields_ = [
        ("fft", KiteOpts),
        ("gerbicz_block", c_u32),
        ("gerbicz_blocks", c_u32),
        ("max_roundoff", c_dbl),
        ("checkpoint", ctypes.c_char_p),
        ("checkpoint_secs", c_dbl),
        ("checkpoint_unverified", ctypes.c_char_p),
        ("unverified_secs", c_dbl),
        ("progress", PROGRESS_FN),
        ("progress_arg", c_p),
        ("inject_fault_at", c_u64),
        ("points", ctypes.POINTER(c_u32)),
        ("npoints", c_u32),
        ("on_point", POINT_FN),
        ("point_arg", c_p),
        ("abort_flag", ctypes.POINTER(ctypes.c_int)),
    ]


class PrpResult(ctypes.Structure):
    _fields_ = [
        ("res64", c_u64),
        ("is_prp", ctypes.c_int),
        ("fft_len", c_u32),
        ("maxerr", c_dbl),
        ("gerbicz_errors", c_u32),
        ("roundoff_retries", c_u32),
        ("resumed", ctypes.c_int),
        ("resumed_at", c_u64),
        ("seconds", c_dbl),
    ]


class LlrOpts(ctypes.Structure):
    _fields_ = [
        ("fft", KiteOpts),
        ("max_roundoff", c_dbl),
        ("checkpoint", ctypes.c_char_p),
        ("checkpoint_secs", c_dbl),
        ("progress", PROGRESS_FN),
        ("on_confirm", CONFIRM_FN),
        ("progress_arg", c_p),
        ("abort_flag", ctypes.POINTER(ctypes.c_int)),
        ("inject_fault_at", c_u64),
        ("inject_fault_runs", c_u32),
        ("inject_roundoff_at", c_u64),
        ("inject_roundoff_ladder", ctypes.c_int),
    ]


class LlrResult(ctypes.Structure):
    _fields_ = [
        ("is_prime", ctypes.c_int),
        ("factor", c_u64),
        ("res64", c_u64),
        ("P", c_u32),
        ("fft_len", c_u32),
        ("maxerr", c_dbl),
        ("roundoff_retries", c_u32),
        ("confirm_runs", c_u32),
        ("resumed", ctypes.c_int),
        ("resumed_at", c_u64),
        ("seconds", c_dbl),
    ]


# name -> (restype, argtypes) for every function declared in include/kite.h
SIGNATURES = {
    "kite_ctx_new": (c_p, [c_u64, c_u64, c_i32, ctypes.POINTER(KiteOpts)]),
    "kite_ctx_free": (None, [c_p]),
    "kite_fft_len": (c_u32, [c_p]),
    "kite_backend_name": (ctypes.c_char_p, [c_p]),
    "kite_num_new": (c_p, [c_p]),
    "kite_num_free": (None, [c_p, c_p]),
    "kite_copy": (None, [c_p, c_p, c_p]),
    "kite_set_u64": (None, [c_p, c_p, c_u64]),
    "kite_square": (None, [c_p, c_p]),
    "kite_square_n": (None, [c_p, c_p, c_u64]),
    "kite_mul": (None, [c_p, c_p, c_p]),
    "kite_mul_small": (None, [c_p, c_p, c_i64]),
    "kite_add": (None, [c_p, c_p, c_p]),
    "kite_sub": (None, [c_p, c_p, c_p]),
    "kite_max_roundoff": (c_dbl, [c_p]),
    "kite_reset_roundoff": (None, [c_p]),
    "kite_set_roundoff_check": (None, [c_p, c_u32]),
    "kite_limbs": (ctypes.c_size_t, [c_p]),
    "kite_get_limbs": (None, [c_p, c_p, ctypes.POINTER(c_u64)]),
    "kite_set_limbs": (None, [c_p, c_p, ctypes.POINTER(c_u64)]),
    "kite_res64": (c_u64, [c_p, c_p]),
    "kite_prp": (ctypes.c_int, [c_u64, c_u64, c_u32, ctypes.POINTER(PrpOpts), ctypes.POINTER(PrpResult)]),
    "kite_llr": (ctypes.c_int, [c_u64, c_u64, ctypes.POINTER(LlrOpts), ctypes.POINTER(LlrResult)]),
    "kite_llr_choose_P": (ctypes.c_int, [c_u64, c_u64, ctypes.POINTER(c_u32), ctypes.POINTER(c_u64)]),
    "kite_llr_checkpoint_iter": (ctypes.c_int, [ctypes.c_char_p, c_u64, c_u64, ctypes.POINTER(c_u64),
                                                ctypes.POINTER(ctypes.c_int)]),
    "kite_prp_checkpoint_iter": (ctypes.c_int, [ctypes.c_char_p, c_u64, c_u64, c_u32, ctypes.POINTER(c_u64)]),
    "kite_debug_corrupt": (None, [c_p, c_p]),
    "kite_choose_fft_len": (c_u32, [c_u64, c_u64]),
    "kite_next_fft_len": (c_u32, [c_u32]),
    "kite_safe_bpw": (c_dbl, [c_u32, c_u64]),
    "kite_residue_nm1": (ctypes.c_int, [ctypes.POINTER(c_u64), ctypes.POINTER(c_u64), ctypes.c_size_t, c_u64]),
    "kite_s_at": (c_u64, [c_p, c_u64]),
    "kite_weight": (c_dbl, [c_p, c_u64, c_u64]),
    "kite_carry_digits": (None, [c_p, ctypes.POINTER(c_i64)]),
}

_lib = None


def lib():
    """Load libkite.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def to_limbs(v, nl):
    arr = (c_u64 * nl)()
    for i in range(nl):
        arr[i] = (v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF
    return arr


def from_limbs(arr):
    return sum(int(x) << (64 * i) for i, x in enumerate(arr))


class Ctx:
    """A kite context for N = k*2^n - 1 on the GPU."""

    def __init__(self, k, n, fft_len=0, plan_order=0):
        self.lib = lib()
        self.k, self.n = k, n
        self.N = k * (1 << n) - 1
        o = KiteOpts(KITE_BACKEND_HIP, fft_len, 0, plan_order)
        self.ptr = self.lib.kite_ctx_new(k, n, -1, ctypes.byref(o))
        if not self.ptr:
            raise RuntimeError(f"kite_ctx_new failed for {k}*2^{n}-1 (fft_len={fft_len})")
        self.L = self.lib.kite_fft_len(self.ptr)
        self.nl = self.lib.kite_limbs(self.ptr)

    def close(self):
        if self.ptr:
            self.lib.kite_ctx_free(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def num(self, value=None):
        p = self.lib.kite_num_new(self.ptr)
        if not p:
            raise RuntimeError("kite_num_new failed")
        x = Num(self, p)
        if value is not None:
            x.set(value)
        return x

    def maxerr(self):
        return self.lib.kite_max_roundoff(self.ptr)

    def reset_maxerr(self):
        self.lib.kite_reset_roundoff(self.ptr)


class Num:
    def __init__(self, ctx, p):
        self.ctx, self.p = ctx, p

    def free(self):
        if self.p:
            self.ctx.lib.kite_num_free(self.ctx.ptr, self.p)
            self.p = None

    def set(self, v):
        arr = to_limbs(v % self.ctx.N, self.ctx.nl)
        self.ctx.lib.kite_set_limbs(self.ctx.ptr, self.p, arr)

    def get(self):
        arr = (c_u64 * self.ctx.nl)()
        self.ctx.lib.kite_get_limbs(self.ctx.ptr, self.p, arr)
        return from_limbs(arr)

    def square(self, count=1):
        self.ctx.lib.kite_square_n(self.ctx.ptr, self.p, count)

    def mul(self, other):
        self.ctx.lib.kite_mul(self.ctx.ptr, self.p, other.p)

    def mul_small(self, m):
        self.ctx.lib.kite_mul_small(self.ctx.ptr, self.p, m)

    def add(self, other):
        self.ctx.lib.kite_add(self.ctx.ptr, self.p, other.p)

    def sub(self, other):
        self.ctx.lib.kite_sub(self.ctx.ptr, self.p, other.p)

    def copy_from(self, other):
        self.ctx.lib.kite_copy(self.ctx.ptr, self.p, other.p)

    def res64(self):
        return self.ctx.lib.kite_res64(self.ctx.ptr, self.p)


def prp(k, n, base=3, block=0, fft_len=0, fault_at=0, checkpoint=None, uckpt=None, abort_flag=None,
        progress=None, ckpt_secs=0.0, uckpt_secs=0.0, max_roundoff=0.0, points=None, on_point=None):
    o = PrpOpts()
    o.fft = KiteOpts(KITE_BACKEND_HIP, fft_len, 0, 0)
    o.gerbicz_block = block
    o.inject_fault_at = fault_at
    o.max_roundoff = max_roundoff
    o.checkpoint = checkpoint.encode() if checkpoint else None
    o.checkpoint_secs = ckpt_secs
    o.checkpoint_unverified = uckpt.encode() if uckpt else None
    o.unverified_secs = uckpt_secs
    keep = []
    if progress is not None:
        cb = PROGRESS_FN(lambda a, i, nn, e: progress(i, nn, e))
        keep.append(cb)
        o.progress = cb
    if abort_flag is not None:
        o.abort_flag = ctypes.pointer(abort_flag)
    if points is not None:
        arr = (c_u32 * len(points))(*points)
        keep.append(arr)
        o.points = arr
        o.npoints = len(points)
        cb2 = POINT_FN(lambda a, idx, it, limbs: on_point(idx, it, limbs))
        keep.append(cb2)
        o.on_point = cb2
    r = PrpResult()
    rc = lib().kite_prp(k, n, base, ctypes.byref(o), ctypes.byref(r))
    return rc, r


def llr(k, n, fft_len=0, **hooks):
    o = LlrOpts()
    o.fft = KiteOpts(KITE_BACKEND_HIP, fft_len, 0, 0)
    for key, val in hooks.items():
        setattr(o, key, val)
    r = LlrResult()
    rc = lib().kite_llr(k, n, ctypes.byref(o), ctypes.byref(r))
    return rc, r
