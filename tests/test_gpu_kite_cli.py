"""The `kite` command line on the GPU: prp, llr, time and throughput print the reference's lines
(tools/kite.c:170-270) with results equal to the reference's vectors; `throughput --batch 1` runs
its tests as one batch."""
import re
import subprocess
import sys

import pytest

from ahnum_amd import kite

pytestmark = pytest.mark.gpu


def run(*args, timeout=240):
    print("kite", " ".join(map(str, args)), flush=True)
    sys.stdout.flush()
    return subprocess.run([kite.BIN_PATH] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)


def test_prp_line_and_expect():
    r = run("prp", 9357, 20000, "--expect", "79291CD2F47CE48C")
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"^9357\*2\^20000-1 is composite  RES64 79291CD2F47CE48C  FFT \d+  [0-9.]+ s \([0-9.]+ us/iter\)"
                     r"  maxerr 0\.\d{4}  gerbicz_errors 0  fft_retries 0$", r.stdout, re.M)
    assert "MATCH expected 79291CD2F47CE48C" in r.stdout
    r = run("prp", "3*2^26459-1")
    assert r.returncode == 0 and "3*2^26459-1 is a probable prime!  RES64 0000000000000001" in r.stdout
    r = run("prp", 9357, 20000, "--expect", "0000000000000000")
    assert r.returncode == 1 and "MISMATCH expected" in r.stdout


def test_prp_fault_caught():
    """CMakeLists.txt:95: prp 1203 90001 --fault 50000 -> 4BE9F1953E350085, with a Gerbicz error."""
    r = run("prp", 1203, 90001, "--fault", 50000, "--expect", "4BE9F1953E350085")
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"gerbicz_errors (\d+)", r.stdout)
    assert m and int(m.group(1)) >= 1


def test_llr_lines(golden):
    rows = {(int(k), int(n)): want for k, n, want in golden("llr_expect.csv")}
    r = run("llr", 3, 41628, "--expect", "prime")
    assert r.returncode == 0 and rows[(3, 41628)] == "prime"
    assert re.search(r"^3\*2\^41628-1 is prime!  P = \d+  FFT \d+  [0-9.]+ s  maxerr 0\.\d{4}  confirm_runs 0"
                     r"  fft_retries 0$", r.stdout, re.M)
    want = rows[(9357, 25000)]
    r = run("llr", 9357, 25000, "--expect", want)
    assert r.returncode == 0 and f"is composite  RES64 {want}" in r.stdout and "confirm_runs 1" in r.stdout


def test_time_line():
    r = run("time", 8045, 659321, "--iters", 2000)
    assert r.returncode == 0
    assert re.search(r"^8045\*2\^659321-1  FFT 49152 \(13\.\d\d bits/word\)  backend \S+$", r.stdout, re.M)
    m = re.search(r"^([0-9.]+) ms/iter  \(([0-9.]+) us\)  maxerr (0\.\d{4})$", r.stdout, re.M)
    assert m and float(m.group(2)) > 0 and float(m.group(3)) < 0.4


@pytest.mark.parametrize("extra,configs", [((), "1x1,4x1"), (("--batch", 1), "16x1,128x1")])
def test_throughput_lines(extra, configs):
    r = run("throughput", 8045, 659321, "--configs", configs, "--secs", 1, *extra)
    assert r.returncode == 0, r.stdout + r.stderr
    rates = {}
    for tok in configs.split(","):
        m = re.search(rf"^{tok}\s+(\d+) sq/s\s+([0-9.]+) tests/hour  \(([0-9.]+) us/iter per test\)$", r.stdout, re.M)
        assert m, r.stdout
        rates[tok] = float(m.group(1))
        assert rates[tok] > 0
    best = max(rates, key=rates.get)
    assert re.search(rf"^best: {best}  [0-9.]+ tests/hour$", r.stdout, re.M)
    if extra:  # a batch of 128 squares more per second than one of 16: the device fills up
        assert rates["128x1"] > 1.5 * rates["16x1"]
