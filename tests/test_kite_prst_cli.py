"""kite-prst, the PRST-compatible command line (reference tools/kite_prst.c): argument parsing, the
hand-off to PRST for everything outside its slice, the proof count check, and -v. These run without
a GPU: the hand-off happens before the engine is touched."""
import os
import stat
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ahnum_amd", "bin", "kite-prst")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="kite-prst not built")


def run(args, env=None, cwd=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([CLI] + args, capture_output=True, text=True, env=e, cwd=cwd, timeout=60)


def fake_prst(tmp_path):
    """A stand-in PRST that records the arguments it was started with."""
    log = tmp_path / "prst_args.txt"
    script = tmp_path / "prst"
    script.write_text(f"#!/bin/sh\nprintf '%s\\n' \"$@\" > {log}\nexit 7\n")
    script.chmod(script.stat().st_mode | stat.S_IEXEC)
    return str(script), log


def test_version():
    r = run(["-v"])
    assert r.returncode == 0
    assert r.stdout.startswith("kite version ")


def test_usage_without_a_number():
    r = run([])
    assert r.returncode == 1
    assert "usage:" in r.stderr


@pytest.mark.parametrize("args", [
    ["9357*2^20000-1", "-d", "-proof", "save", "512"],                       # a proof of an LLR run
    ["9357*2^20000-1", "-d", "-fermat", "a", "3", "-proof", "merge", "512"],  # no such proof step here
    ["9357*2^20000+1", "-d"],                                                 # not k*2^n-1
    ["9358*2^20000-1", "-d", "-fermat", "a", "3"],                            # k even
    ["9357*2^20000-1", "-d", "-fermat", "a", "3", "-oddoption"],
    ["9357*2^20000-1", "-fft", "info"],                                       # -fft info of an LLR run
])
def test_outside_the_slice_goes_to_prst(tmp_path, args):
    """tools/kite_prst.c:98-182: anything it does not run is handed to PRST with the same arguments
    ($KITE_PRST), and PRST's exit status is kite-prst's."""
    prst, log = fake_prst(tmp_path)
    r = run(args, env={"KITE_PRST": prst})
    assert r.returncode == 7
    assert log.read_text().split("\n")[:-1] == args


def test_missing_prst_is_an_error(tmp_path):
    r = run(["9357*2^20000+1", "-d"], env={"KITE_PRST": str(tmp_path / "nowhere")})
    assert r.returncode == 2
    assert "cannot run PRST" in r.stderr


def test_proof_count_must_be_a_power_of_two():
    r = run(["9357*2^20000-1", "-d", "-fermat", "a", "3", "-proof", "save", "384"])
    assert r.returncode == 1
    assert "proof count is not a power of 2." in r.stderr
