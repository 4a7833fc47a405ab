"""ctypes binding to the MI355X kite library (ahnum_amd/lib/libkite.so).

Mirrors the reference C API (include/kite.h) one to one for tests and the benchmark. The library
is the product: there is no Python or CPU fallback, and loading fails loudly when it is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("KITE_LIB") or os.path.join(HERE, "lib", "libkite.so")  # KITE_LIB: another build of the library
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
    """The low 64*nl bits of v >= 0 as nl little-endian 64-bit limbs (linear in the size)."""
    v &= (1 << (64 * nl)) - 1
    return (c_u64 * nl).from_buffer_copy(v.to_bytes(8 * nl, "little"))


def from_limbs(arr):
    return int.from_bytes(bytes(arr), "little")


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


# ---- batched engine (include/kite_batch.h) -----------------------------------------------------

class BatchTest(ctypes.Structure):
    _fields_ = [("k", c_u64), ("n", c_u64), ("base", c_u32), ("fft_len", c_u32)]


DONE_FN = ctypes.CFUNCTYPE(None, c_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(PrpResult))


class BatchPrpOpts(ctypes.Structure):
    _fields_ = [
        ("slots", c_u32),
        ("gerbicz_block", c_u32),
        ("gerbicz_blocks", c_u32),
        ("max_roundoff", c_dbl),
        ("inject_fault_test", c_u64),
        ("inject_fault_at", c_u64),
        ("on_done", DONE_FN),
        ("arg", c_p),
    ]


I32P = ctypes.POINTER(ctypes.c_int32)
U32P = ctypes.POINTER(c_u32)
SIGNATURES.update({
    "kite_batch_new": (c_p, []),
    "kite_batch_free": (None, [c_p]),
    "kite_batch_add": (ctypes.c_int, [c_p, c_u64, c_u64, c_u32, c_u32]),
    "kite_batch_remove": (None, [c_p, ctypes.c_int]),
    "kite_batch_capacity": (ctypes.c_int, [c_p]),
    "kite_batch_ctx": (c_p, [c_p, ctypes.c_int]),
    "kite_batch_num": (c_p, [c_p, ctypes.c_int, c_u32]),
    "kite_batch_square": (ctypes.c_int, [c_p, I32P, U32P]),
    "kite_batch_mul": (ctypes.c_int, [c_p, I32P, I32P]),
    "kite_batch_copy": (ctypes.c_int, [c_p, I32P, I32P]),
    "kite_batch_equal": (ctypes.c_int, [c_p, I32P, I32P, I32P]),
    "kite_batch_maxerr": (ctypes.c_int, [c_p, ctypes.POINTER(c_dbl), ctypes.c_int]),
    "kite_batch_sync": (ctypes.c_int, [c_p]),
    "kite_batch_set_timing": (ctypes.c_int, [c_p, ctypes.c_int]),
    "kite_batch_pass_times": (ctypes.c_int, [c_p, ctypes.POINTER(c_dbl), ctypes.POINTER(c_u64)]),
    "kite_batch_kernel_times": (ctypes.c_int, [c_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                               ctypes.POINTER(c_dbl), ctypes.POINTER(c_u64),
                                               ctypes.POINTER(c_dbl)]),
    "kite_batch_prp": (ctypes.c_int, [ctypes.POINTER(BatchTest), ctypes.c_size_t, ctypes.POINTER(BatchPrpOpts),
                                      ctypes.POINTER(PrpResult), ctypes.POINTER(ctypes.c_int)]),
})


JOB_PRP, JOB_LLR, JOB_PENDING = 0, 1, 1
ORDER_ARRIVAL, ORDER_LEAST_WORK = 0, 1
SHARE_SQUARINGS, SHARE_TIME = 0, 1  # kite_run_opts.share
LAUNCH_GRAPH, LAUNCH_DIRECT = 0, 1  # kite_run_opts.launch


class Job(ctypes.Structure):
    _fields_ = [("k", c_u64), ("n", c_u64), ("kind", c_u32), ("base", c_u32), ("fft_len", c_u32),
                ("arrival", c_dbl)]


class JobResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("is_prime", ctypes.c_int),
        ("res64", c_u64),
        ("factor", c_u64),
        ("P", c_u32),
        ("fft_len", c_u32),
        ("maxerr", c_dbl),
        ("gerbicz_errors", c_u32),
        ("roundoff_retries", c_u32),
        ("confirm_runs", c_u32),
        ("start", c_dbl),
        ("finish", c_dbl),
        ("resumed_at", c_u64),
        ("resumed", c_u32),
        ("run", c_u32),
        ("iter", c_u64),
        ("verified", c_u64),
    ]


JOB_DONE_FN = ctypes.CFUNCTYPE(None, c_p, ctypes.c_size_t, ctypes.POINTER(JobResult))


class RunOpts(ctypes.Structure):
    _fields_ = [
        ("slots", c_u32),
        ("hbm_budget", c_u64),
        ("block", c_u32),
        ("gerbicz_blocks", c_u32),
        ("max_roundoff", c_dbl),
        ("horizon", c_dbl),
        ("inject_fault_job", c_u64),
        ("inject_fault_at", c_u64),
        ("on_done", JOB_DONE_FN),
        ("arg", c_p),
        ("checkpoint_dir", ctypes.c_char_p),
        ("checkpoint_secs", c_dbl),
        ("order", c_u32),
        ("share", c_u32),
        ("parts", c_u32),
        ("shape_parts", c_u32),
        ("xcds", c_u32),
        ("launch", c_u32),
    ]


SIGNATURES.update({
    "kite_batch_set_xadd": (ctypes.c_int, [c_p, ctypes.c_int, ctypes.c_int32]),
    "kite_batch_set_graphs": (ctypes.c_int, [c_p, ctypes.c_int]),
    "kite_batch_set_parts": (ctypes.c_int, [c_p, ctypes.c_int]),
    "kite_batch_set_xcds": (ctypes.c_int, [c_p, ctypes.c_int]),
    "kite_batch_set_persistent": (ctypes.c_int, [c_p, ctypes.c_int]),
    "kite_batch_set_shape_parts": (ctypes.c_int, [c_p, ctypes.c_int]),
    "kite_debug_fail_hip": (None, [ctypes.c_long]),
    "kite_job_bytes": (c_u64, [c_u32, c_u32]),
    "kite_batch_run": (ctypes.c_int, [ctypes.POINTER(Job), ctypes.c_size_t, ctypes.POINTER(RunOpts),
                                      ctypes.POINTER(JobResult)]),
})


def batch_run(jobs, slots=0, hbm_budget=0, block=1000, blocks=0, horizon=0.0, fault_job=0, fault_at=0,
              max_roundoff=0.0, on_done=None, checkpoint_dir=None, checkpoint_secs=0.0, order=ORDER_ARRIVAL,
              share=SHARE_SQUARINGS, parts=0, shape_parts=0, xcds=0, launch=0):
    """Runs jobs [(kind, k, n, arrival), ...] (kind JOB_PRP or JOB_LLR; Fermat base 3) through
    kite_batch_run, with checkpoint files in checkpoint_dir if given. Returns (rc, [JobResult])."""
    arr = (Job * len(jobs))(*[Job(k, n, kind, 3, 0, arrival) for kind, k, n, arrival in jobs])
    o = RunOpts()
    o.slots = slots
    o.hbm_budget = hbm_budget
    o.block = block
    o.gerbicz_blocks = blocks
    o.max_roundoff = max_roundoff
    o.horizon = horizon
    o.inject_fault_job = fault_job
    o.inject_fault_at = fault_at
    o.order = order
    o.share = share
    o.parts = parts
    o.shape_parts = shape_parts
    o.xcds = xcds
    o.launch = launch
    if checkpoint_dir is not None:
        o.checkpoint_dir = os.fsencode(checkpoint_dir)
        o.checkpoint_secs = checkpoint_secs
    keep = None
    if on_done is not None:
        keep = JOB_DONE_FN(lambda a, i, r: on_done(i, r.contents))
        o.on_done = keep
    res = (JobResult * len(jobs))()
    rc = lib().kite_batch_run(arr, len(jobs), ctypes.byref(o), res)
    return rc, [res[i] for i in range(len(jobs))]


def batch_prp(tests, slots=0, block=1000, blocks=0, fault_test=0, fault_at=0, on_done=None):
    """Runs Fermat base-3 tests [(k, n), ...] through the batched engine. Returns [(status, PrpResult)]."""
    arr = (BatchTest * len(tests))(*[BatchTest(k, n, 3, 0) for k, n in tests])
    o = BatchPrpOpts()
    o.slots = slots
    o.gerbicz_block = block
    o.gerbicz_blocks = blocks
    o.inject_fault_test = fault_test
    o.inject_fault_at = fault_at
    keep = None
    if on_done is not None:
        keep = DONE_FN(lambda a, i, st, r: on_done(i, st, r.contents))
        o.on_done = keep
    res = (PrpResult * len(tests))()
    st = (ctypes.c_int * len(tests))()
    rc = lib().kite_batch_prp(arr, len(tests), ctypes.byref(o), res, st)
    return rc, [(st[i], res[i]) for i in range(len(tests))]


class Batch:
    """A kite_batch: slots of independent residues advanced by shared launches."""

    def __init__(self):
        self.lib = lib()
        self.p = self.lib.kite_batch_new()
        if not self.p:
            raise RuntimeError("kite_batch_new failed (no GPU?)")
        self.slots = {}

    def add(self, k, n, fft_len=0, nnums=1):
        s = self.lib.kite_batch_add(self.p, k, n, fft_len, nnums)
        if s < 0:
            raise RuntimeError(f"kite_batch_add failed for {k}*2^{n}-1")
        self.slots[s] = (k, n)
        return s

    def ctx(self, s):
        return self.lib.kite_batch_ctx(self.p, s)

    def set(self, s, i, v):
        c = self.ctx(s)
        nl = self.lib.kite_limbs(c)
        if self.lib.kite_batch_sync(self.p):
            raise RuntimeError("kite batch: a HIP call failed")
        self.lib.kite_set_limbs(c, self.lib.kite_batch_num(self.p, s, i), to_limbs(v, nl))

    def get(self, s, i):
        c = self.ctx(s)
        nl = self.lib.kite_limbs(c)
        if self.lib.kite_batch_sync(self.p):
            raise RuntimeError("kite batch: a HIP call failed")
        arr = (c_u64 * nl)()
        self.lib.kite_get_limbs(c, self.lib.kite_batch_num(self.p, s, i), arr)
        return from_limbs(arr)

    def fft_len(self, s):
        return self.lib.kite_fft_len(self.ctx(s))

    def square(self, count, which=0):
        cap = self.lib.kite_batch_capacity(self.p)
        w = (ctypes.c_int32 * cap)(*[which if s in self.slots else -1 for s in range(cap)])
        c = (c_u32 * cap)(*[count] * cap)
        if self.lib.kite_batch_square(self.p, w, c):
            raise RuntimeError("kite_batch_square failed")

    def maxerr(self, reset=False):
        cap = self.lib.kite_batch_capacity(self.p)
        out = (c_dbl * cap)()
        self.lib.kite_batch_maxerr(self.p, out, int(reset))
        return {s: out[s] for s in self.slots}

    def sync(self):
        if self.lib.kite_batch_sync(self.p):
            raise RuntimeError("kite batch: a HIP call failed")

    def parts(self, n):
        """Deal the slots to n concurrent streams (1..4); returns the count in use."""
        return self.lib.kite_batch_set_parts(self.p, int(n))

    def xcds(self, n):
        """Lay column jobs out for n XCDs (each test's column groups on one XCD; 1: slot order)."""
        return self.lib.kite_batch_set_xcds(self.p, int(n))

    def graphs(self, on):
        """Replay multi-step square calls from a captured HIP graph (the default) or launch directly
        (on streams with hardware queues of their own); switching waits for the batch's work."""
        if self.lib.kite_batch_set_graphs(self.p, int(bool(on))):
            raise RuntimeError("kite_batch_set_graphs failed (a HIP call failed)")

    def shape_parts(self, on):
        """Each FFT shape but the most common on a stream of its own (kite_batch_set_shape_parts)."""
        return self.lib.kite_batch_set_shape_parts(self.p, int(bool(on)))

    def persistent(self, on):
        """Persistent launches of the square-mode passes (kite_batch_set_persistent)."""
        return self.lib.kite_batch_set_persistent(self.p, int(bool(on)))

    def timing(self, every):
        """Time each launch of every `every`-th step of later square calls (0/False: off)."""
        self.lib.kite_batch_set_timing(self.p, int(every))

    def pass_times(self):
        ms = (c_dbl * 3)()
        n = (c_u64 * 3)()
        self.lib.kite_batch_pass_times(self.p, ms, n)
        return list(ms), list(n)

    def kernel_times(self):
        """[(name, total_ms, launches, algorithmic_bytes)] of the last timed square call."""
        out = []
        cnt = self.lib.kite_batch_kernel_times(self.p, -1, None, 0, None, None, None)
        for i in range(cnt):
            name = ctypes.create_string_buffer(64)
            ms, ln, by = c_dbl(), c_u64(), c_dbl()
            self.lib.kite_batch_kernel_times(self.p, i, name, 64, ctypes.byref(ms), ctypes.byref(ln),
                                             ctypes.byref(by))
            if ln.value:
                out.append((name.value.decode(), ms.value, ln.value, by.value))
        return out

    def close(self):
        if self.p:
            self.lib.kite_batch_free(self.p)
            self.p = None
