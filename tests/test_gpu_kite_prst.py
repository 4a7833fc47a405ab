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


# tests/prst_compat.sh: PRST 14.0.1888's files for 1203*2^90001-1 -d -fermat a 3 -proof save 512
# keep, then -proof build 512, then -proof cert (MD5s of the files; RES64s of the lines)
PRST_MD5 = {
    "prst_1107603830.proof.0": "dafcb9cf37507263eaee0ece52991b3e",
    "prst_1107603830.proof.512": "c57ae2618fb59c8652987543ab7cec29",
    "prst_1107603830.prod.0": "57b767aa53d89f86a5fe742d66089abe",
    "prst_1107603830.prod.1": "2d426d68ba20d327d66d58a1e0c33316",
    "prst_1107603830.prod.2": "c3b3acd901f61818ad42b61337cadf9b",
    "prst_1107603830.prod.3": "bfcbaed0176cd4a985744b04692d21f8",
    "prst_1107603830.prod.4": "a58741b2b37270b7f213119e48d9ae4a",
    "prst_1107603830.prod.5": "20d54b4ed93bd466f69c39a51a27113b",
    "prst_1107603830.prod.6": "b8f2e1c7af03a5d277007f944335876e",
    "prst_1107603830.prod.7": "1d8bcc840e8d7c3246b17898d1b96b34",
    "prst_1107603830.prod.8": "3d38d7470434544fa95858a251bd3b8a",
}
PRST_CERT_MD5 = "f3ae012634454e38aa7311deccccde56"


def md5_of(path):
    import hashlib
    return hashlib.md5(open(path, "rb").read()).hexdigest()


def test_proof_files_match_prst_byte_for_byte(tmp_path):
    """Proof points, products and certificate identical to PRST's (by MD5), the raw and final
    certificate RES64s PRST prints, and the certificate verifying in a fresh directory."""
    expr = "1203*2^90001-1"
    r, out = run([expr, "-d", "-fermat", "a", "3", "-proof", "save", "512", "keep"], tmp_path, timeout=600)
    assert r.returncode == 0, r.stderr
    assert f"{expr} Saving 512 proof points." in out
    assert "is not prime. RES64: 4BE9F1953E350085" in out
    assert re.search(rf"{re.escape(expr)} compressed 512 points to 9 products, time: [0-9.]+ s\.", out)
    assert "raw certificate RES64: EC81DC6893506908" in out
    got = {f: md5_of(tmp_path / f) for f in PRST_MD5}
    assert got == PRST_MD5
    for f in PRST_MD5:  # every file carries its MD5 sidecar
        assert (tmp_path / (f + ".md5")).read_text() == PRST_MD5[f]
    assert not [f for f in os.listdir(tmp_path) if f.startswith("kite_") and f.endswith(".done")]

    r, out = run([expr, "-d", "-fermat", "a", "3", "-proof", "build", "512"], tmp_path, timeout=600)
    assert r.returncode == 0, r.stderr
    assert f"{expr} Building certificate from 9 products." in out
    assert "certificate RES64: EC81DC6893506908" in out
    assert md5_of(tmp_path / "prst_1107603830.cert") == PRST_CERT_MD5
    assert not (tmp_path / "prst_1107603830.proof.0").exists()  # removed without keep

    v = tmp_path / "v"
    v.mkdir()
    (v / "verify.cert").write_bytes((tmp_path / "prst_1107603830.cert").read_bytes())
    r, out = run([expr, "-d", "-proof", "cert", "verify.cert"], v, timeout=600)
    assert r.returncode == 0, r.stderr
    assert re.search(rf"Verifying certificate of {re.escape(expr)}, complexity = \d+\.", out)
    assert "certificate RES64: EC81DC6893506908" in out
    lines = (tmp_path / "result.txt").read_text().splitlines()
    assert any("raw certificate RES64: EC81DC6893506908" in x for x in lines)
