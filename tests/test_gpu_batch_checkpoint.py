"""Checkpoints of kite_batch_run: jobs that the horizon stops leave their verified state in the
checkpoint directory (the KITECKP1 / KITELLR1 formats of the reference's src/prp.c:156-212 and
src/llr.c:92-146, under kite-prst's file names); the next run resumes them and ends with the known
results, and kite-prst finishes a test that a batch stopped."""
import os
import re
import struct
import subprocess

import pytest

from ahnum_amd import kite

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ahnum_amd", "bin", "kite-prst")
FINGERPRINT_P = 3417905339  # files are named by N mod this prime

FERMAT = (9357, 369717, "F5FDA01218F20439")  # tests/known_residues.csv
LLR_PRIME = (9001, 662695)                   # tests/benchmarks.csv:2, a known prime


def say(*a):
    print(*a, flush=True)


def fingerprint(k, n):
    return (k * pow(2, n, FINGERPRINT_P) - 1) % FINGERPRINT_P


def fermat_file(d, k, n, base=3):
    return d / f"kite_{fingerprint(k, n)}.{base}.ckpt"


def llr_file(d, k, n):
    return d / f"kite_{fingerprint(k, n)}.llr.ckpt"


def read_fermat(path):
    raw = path.read_bytes()
    magic = raw[:8]
    k, n, base, it, nl, L = struct.unpack("<6Q", raw[8:56])
    x, d = raw[56:56 + 8 * nl], raw[56 + 8 * nl:56 + 16 * nl]
    assert len(raw) == 56 + 16 * nl + 8
    return magic, (k, n, base), it, L, x, d


def read_llr(path):
    raw = path.read_bytes()
    magic = raw[:8]
    k, n, P, it, nl, L, run, p0, p1 = struct.unpack("<9Q", raw[8:80])
    assert len(raw) == 80 + 8 * nl + 8
    return magic, (k, n, P), it, L, run


def stop_at_horizon(d, extra=()):
    """Fermat and LLR jobs stopped by a 3 s horizon (Gerbicz checks every 8 blocks of 1000)."""
    fk, fn, _ = FERMAT
    lk, ln = LLR_PRIME
    jobs = [(kite.JOB_PRP, fk, fn, 0.0), (kite.JOB_LLR, lk, ln, 0.0)] + list(extra)
    say("batch to the horizon:", jobs)
    rc, out = kite.batch_run(jobs, block=1000, blocks=8, horizon=3.0, checkpoint_dir=str(d), checkpoint_secs=0.5)
    for j, r in zip(jobs, out):
        say("  ", j, "status", r.status, "fft", r.fft_len)
    assert out[0].status == kite.JOB_PENDING and out[1].status == kite.JOB_PENDING
    magic, key, it, L, x, dd = read_fermat(fermat_file(d, fk, fn))
    assert magic == b"KITECKP1" and key == (fk, fn, 3)
    assert 0 < it < fn and it % 1000 == 0 and L == out[0].fft_len
    assert x == dd  # written with d = x: a reader restarts the Gerbicz product at x
    magic, key, lit, lL, run = read_llr(llr_file(d, lk, ln))
    assert magic == b"KITELLR1" and key == (lk, ln, out[1].P) and run == 0 and 0 < lit < ln - 1
    say("  files at Fermat iteration", it, "LLR step", lit)
    return out, it, lit


def test_horizon_checkpoint_then_resume(tmp_path):
    """Run to a horizon, then again: each job resumes where its file says and ends with the known
    result; a job that finished leaves no file, and the files are gone once all finish."""
    fk, fn, want = FERMAT
    lk, ln = LLR_PRIME
    out, it, lit = stop_at_horizon(tmp_path, extra=[(kite.JOB_PRP, 9357, 20000, 0.0)])
    assert out[2].status == 0 and "%016X" % out[2].res64 == "79291CD2F47CE48C"
    assert not out[2].resumed and not fermat_file(tmp_path, 9357, 20000).exists()
    jobs = [(kite.JOB_PRP, fk, fn, 0.0), (kite.JOB_LLR, lk, ln, 0.0)]
    say("batch resumed:", jobs)
    rc, out = kite.batch_run(jobs, block=1000, checkpoint_dir=str(tmp_path))
    assert rc == 0
    assert out[0].resumed and out[0].resumed_at == it
    assert "%016X" % out[0].res64 == want and not out[0].is_prime and out[0].maxerr < 0.4
    assert out[1].resumed and out[1].resumed_at == lit
    assert out[1].is_prime and out[1].confirm_runs == 0 and out[1].maxerr < 0.4
    assert not [f for f in os.listdir(tmp_path) if f.endswith(".ckpt")]


def test_kite_prst_finishes_what_a_batch_stopped(tmp_path):
    """The batch's files under kite-prst's names: kite-prst announces the resume and ends with the
    same RES64 (Fermat, restarting its Gerbicz product at the file's x) and the prime (LLR)."""
    fk, fn, want = FERMAT
    lk, ln = LLR_PRIME
    stop_at_horizon(tmp_path)
    for args, expect in [([f"{fk}*2^{fn}-1", "-d", "-fermat", "a", "3"], f"is not prime. RES64: {want}"),
                         ([f"{lk}*2^{ln}-1", "-d"], "is prime!")]:
        say("kite-prst", " ".join(args))
        r = subprocess.run([CLI] + args, capture_output=True, text=True, cwd=tmp_path, timeout=300)
        out = r.stdout.replace("\r", "\n")
        assert r.returncode == 0, r.stderr
        m = re.search(r"restarting at (\d+\.\d)%\.", out)
        assert m and float(m.group(1)) > 0, out[-400:]
        assert expect in out, out[-400:]
    assert not [f for f in os.listdir(tmp_path) if f.endswith(".ckpt")]


def test_record_size_gerbicz_checked_squarings(tmp_path):
    """The record-size configuration: 91*2^24767982-1 (workloads.record, FFT 1638400, the reference's
    length for it, README.md:202) runs for a
    20 s horizon with a Gerbicz check every 10 blocks of 1000; at least 10^5 squarings are verified
    (the checkpoint's iteration) with no check failing and round-off below 0.4."""
    from ahnum_amd import workloads
    (k, n), = workloads.record()
    say("record", k, n)
    rc, out = kite.batch_run([(kite.JOB_PRP, k, n, 0.0)], block=1000, blocks=10, horizon=20.0,
                             checkpoint_dir=str(tmp_path))
    r = out[0]
    say("status", r.status, "fft", r.fft_len, "maxerr", r.maxerr, "gerbicz errors", r.gerbicz_errors)
    assert r.status == kite.JOB_PENDING and r.fft_len == 1638400
    assert r.gerbicz_errors == 0 and r.roundoff_retries == 0 and 0 < r.maxerr < 0.4
    magic, key, it, L, x, d = read_fermat(fermat_file(tmp_path, k, n))
    say("verified iterations", it)
    assert magic == b"KITECKP1" and key == (k, n, 3) and L == 1638400 and it >= 100000 and it % 10000 == 0


@pytest.mark.skipif(not os.environ.get("KITE_LONG_TESTS"), reason="about 6 minutes of GPU: set KITE_LONG_TESTS=1")
def test_large_fft_reference_residues_in_checkpointed_chunks(tmp_path, golden):
    """The reference's tests/large_fft_cases.csv: 7 numbers at n 2.3M-5.5M (FFT 160K-442K) whose
    RES64s come from PRST and kite, in one batch run as 60 s horizon chunks that resume from their
    checkpoint files (Gerbicz check every 20 blocks of 1000). Every RES64 must match."""
    rows = golden("large_fft_cases.csv")
    pending = {i: (int(k), int(n)) for i, (k, n, *_rest) in enumerate(rows)}
    results = {}
    for chunk in range(15):
        if not pending:
            break
        idx = sorted(pending)
        rc, out = kite.batch_run([(kite.JOB_PRP, *pending[i], 0.0) for i in idx], block=1000, blocks=20,
                                 horizon=60.0, checkpoint_dir=str(tmp_path))
        for i, r in zip(idx, out):
            if r.status != kite.JOB_PENDING:
                results[i] = r
                del pending[i]
        say("chunk", chunk, "finished", sorted(results), "resumed", [r.resumed_at for r in out])
    assert not pending
    for i, (k, n, want, ref_fft, _src) in enumerate(rows):
        r = results[i]
        say(k, n, r.status, "%016X" % r.res64, "fft", r.fft_len, "(reference", ref_fft + ")", "maxerr", r.maxerr)
        assert r.status == 0 and "%016X" % r.res64 == want and r.maxerr < 0.4, (k, n)
