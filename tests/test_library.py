"""CPU checks of the C-ABI library: it loads, exports every function include/*.h declares, and
its host-only functions (length choice, residue conversion, P choice) agree with the oracle and
exact arithmetic. No GPU work is launched here."""
import ctypes
import os
import re
import subprocess

import pytest

from ahnum_amd import kite
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
        for m in re.finditer(r"\b(kite\w*)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = kite.lib()
    names = declared_functions()
    assert len(names) >= 34
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_nm_exports_match_header():
    out = subprocess.run(["nm", "-D", "--defined-only", kite.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in declared_functions():
        assert n in exported, n


def test_fft_length_choice_matches_reference(golden):
    """With compile-time column kernels for odd parts 1, 3, 5, 7, 9, 15, 25 and 27, and no tuning
    file, the selector makes the reference's choice for its 7 production cases
    (tests/large_fft_cases.csv:4-10, among them 294912 = 2^15 9 and 442368 = 2^14 27) and for the
    record test (1638400 = 2^16 25, README.md:202)."""
    lib = kite.lib()
    for k, n, _, Lref, _ in golden("large_fft_cases.csv"):
        k, n, Lref = int(k), int(n), int(Lref)
        L = lib.kite_choose_fft_len(k, n)
        assert L == Lref and n / L <= lib.kite_safe_bpw(L, k), (k, n)
    assert lib.kite_choose_fft_len(9001, 662695) == 49152
    assert lib.kite_choose_fft_len(91, 24767982) == 1638400


def test_next_fft_len():
    lib = kite.lib()
    for L in [4096, 49152, 5000, 163840]:
        M = lib.kite_next_fft_len(L)
        assert M > L and M % 64 == 0 and O.lib().or_fft_len_ok(M)


def test_residue_nm1_matches_exact():
    lib = kite.lib()
    for k, n in [(9357, 200), (8191, 250), (5, 301), (3, 100), (1, 127), (1203, 611)]:
        N = k * (1 << n) - 1
        nl = (n + 127) // 64
        y = kite.to_limbs(pow(3, N + 1, N), nl)
        assert lib.kite_residue_nm1(y, kite.to_limbs(N, nl), nl, 3) == 0
        assert kite.from_limbs(y) == pow(3, N - 1, N)


def test_llr_choose_P_matches_oracle():
    lib = kite.lib()
    for k, n in [(3, 26459), (9357, 20000), (33, 99999), (4299, 522773), (5, 1000), (65, 3001)]:
        P1, f1 = ctypes.c_uint32(), ctypes.c_uint64()
        P2, f2 = ctypes.c_uint32(), ctypes.c_uint64()
        assert lib.kite_llr_choose_P(k, n, ctypes.byref(P1), ctypes.byref(f1)) == 0
        assert O.lib().or_llr_choose_P(k, n, ctypes.byref(P2), ctypes.byref(f2)) == 0
        assert (P1.value, f1.value) == (P2.value, f2.value)


def test_cli_usage_without_gpu():
    r = subprocess.run([kite.BIN_PATH], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_ctx_rejects_cpu_backends_and_bad_input():
    lib = kite.lib()
    for backend in (1, 2, 3, 4):  # REF, NEON, SME, C2P are not built
        o = kite.KiteOpts(backend, 0, 0, 0)
        assert not lib.kite_ctx_new(9357, 20000, -1, ctypes.byref(o))
    o = kite.KiteOpts(0, 0, 0, 0)
    assert not lib.kite_ctx_new(9358, 20000, -1, ctypes.byref(o))  # even k
    assert not lib.kite_ctx_new(9357, 20000, 1, ctypes.byref(o))   # c != -1


def test_batch_run_refuses_an_unusable_checkpoint_dir(tmp_path):
    """kite_batch_run with a checkpoint directory that is missing, or a file, returns -1 at once
    with every status -1, before it touches the GPU (include/kite_batch.h)."""
    f = tmp_path / "a_file"
    f.write_text("x")
    for d in (tmp_path / "missing", f):
        rc, out = kite.batch_run([(kite.JOB_PRP, 9357, 20000, 0.0), (kite.JOB_LLR, 3, 26459, 0.0)],
                                 checkpoint_dir=str(d))
        assert rc == -1 and [r.status for r in out] == [-1, -1]


def test_bpw_fit_recovers_model_constants():
    """The least-squares fit of ahnum_amd/fit_bpw.py returns the constants of limits generated from
    the model itself (so a sign or column mix-up in the design matrix fails here)."""
    from ahnum_amd import fit_bpw
    import math
    rows = [(L, k, 21.0 - 0.3 * (math.log2(L) - 12) - 0.55 * math.log2(k))
            for L in (4096, 49152, 1048576) for k in (1, 3, 255, 9999)]
    (A, B, C), rms = fit_bpw.fit(rows)
    assert abs(A - 21.0) < 1e-9 and abs(B - 0.3) < 1e-9 and abs(C - 0.55) < 1e-9 and rms < 1e-9
