"""kite-prst runs on the GPU: the PRST-format stdout lines and result.txt (the reference's
docs/prst-compat.md line list), RES64s against GMP and the reference's vectors, LLR verdicts, and
SIGTERM -> checkpoint -> resume (the reference's tests/prst_kill.py behaviour)."""
import os
import re
import signal
import subprocess
import sys
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ahnum_amd", "bin", "kite-prst")

pytestmark = pytest.mark.gpu

FINGERPRINT_P = 3417905339  # PRST names its files by N mod this prime


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(args, cwd, timeout=240):
    say("kite-prst", " ".join(args))
    r = subprocess.run([CLI] + args, capture_output=True, text=True, cwd=cwd, timeout=timeout)
    return r, r.stdout.replace("\r", "\n")


def test_fermat_composite_lines_and_result_file(tmp_path):
    r, out = run(["9357*2^20000-1", "-d", "-fermat", "a", "3"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert re.search(r"Using (.*?)\s*FFT length (\d+)([KM]?)", out)
    m = re.search(r"Fermat probabilistic test of 9357\*2\^20000-1, a = (\d+), complexity = (\d+)\.", out)
    assert m and m.group(1) == "3" and int(m.group(2)) == 20000 + 20000 // 141
    assert "9357*2^20000-1 Gerbicz check enabled, L2 = 141*35." in out
    assert re.search(r"9357\*2\^20000-1 is not prime\. RES64: 79291CD2F47CE48C, time: [0-9.]+ s\.", out)
    res = (tmp_path / "result.txt").read_text()
    assert re.fullmatch(r"9357\*2\^20000-1 is not prime\. RES64: 79291CD2F47CE48C, time: \d+ s\.\n", res)
    assert not [f for f in os.listdir(tmp_path) if f.endswith(".ckpt")]  # removed on completion


def test_fermat_probable_prime(tmp_path):
    r, out = run(["3*2^26459-1", "-d", "-fermat", "a", "3"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert re.search(r"3\*2\^26459-1 is a probable prime\. Time: [0-9.]+ s\.", out)


def test_fermat_divisible_by_base(tmp_path):
    # 5*2^1001 - 1 = 9 = 0 (mod 3)
    r, out = run(["5*2^1001-1", "-d", "-fermat", "a", "3"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "5*2^1001-1 is not prime, divisible by 3." in out
    assert "RES64" not in out


def test_fft_info_stops_after_the_test_line(tmp_path):
    r, out = run(["9357*2^20000-1", "-fermat", "a", "3", "-fft", "info"], tmp_path)
    assert r.returncode == 0
    assert "Fermat probabilistic test of" in out and "Gerbicz" not in out


def test_llr_prime_and_composite(tmp_path, golden):
    rows = {(int(k), int(n)): want for k, n, want in golden("llr_expect.csv")}
    r, out = run(["3*2^41628-1", "-d"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert re.search(r"Morrison \(LLR\) test of 3\*2\^41628-1, P = \d+, Q = -1, complexity = 41628\.", out)
    assert rows[(3, 41628)] == "prime" and re.search(r"3\*2\^41628-1 is prime! Time: [0-9.]+ s\.", out)
    r, out = run(["9357*2^20000-1", "-d"], tmp_path)
    assert r.returncode == 0, r.stderr
    want = rows[(9357, 20000)]
    assert f"9357*2^20000-1 is not a probable prime. Have you run Fermat test first? RES64: {want}, time:" in out
    assert "is prime!" in open(tmp_path / "result.txt").read()


def test_sigterm_checkpoint_and_resume(tmp_path):
    """tests/prst_kill.py: a SIGTERM mid-test writes the checkpoint, prints Terminating... and exits 1;
    the same command resumes ("restarting at P%.") and ends with the reference's RES64."""
    k, n = 9357, 200001
    expr = f"{k}*2^{n}-1"
    args = [expr, "-d", "-fermat", "a", "3", "-time", "write", "1", "progress", "1"]
    say("kite-prst (to be stopped)", " ".join(args))
    p = subprocess.Popen([CLI] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=tmp_path)
    ckpt_u = tmp_path / f"kite_{(k * pow(2, n, FINGERPRINT_P) - 1) % FINGERPRINT_P}.3.u.ckpt"
    t0 = time.time()
    while not ckpt_u.exists() and time.time() - t0 < 120 and p.poll() is None:
        time.sleep(0.2)
    assert ckpt_u.exists(), "no unverified checkpoint appeared"
    p.send_signal(signal.SIGTERM)
    out, err = p.communicate(timeout=120)
    say("stopped:", out.replace("\r", "\n").strip().split("\n")[-1])
    assert p.returncode == 1 and "Terminating..." in out
    r, out = run(args, tmp_path)
    assert r.returncode == 0, r.stderr
    m = re.search(r"restarting at (\d+\.\d)%\.", out)
    assert m and float(m.group(1)) > 0
    assert f"{expr} is not prime. RES64: BDD48C45E17CDD83" in out
