"""The ctypes mirrors in ahnum_amd/kite.py have exactly the C layouts of include/kite.h and
include/kite_batch.h: the size of every struct and the offset of every field, as the C compiler
lays them out (a mismatch would corrupt every call that passes one). No GPU."""
import ctypes
import os
import subprocess

import pytest

from ahnum_amd import kite

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [  # (C type, ctypes mirror)
    ("kite_opts", kite.KiteOpts),
    ("kite_prp_opts", kite.PrpOpts),
    ("kite_prp_result", kite.PrpResult),
    ("kite_llr_opts", kite.LlrOpts),
    ("kite_llr_result", kite.LlrResult),
    ("kite_batch_test", kite.BatchTest),
    ("kite_batch_prp_opts", kite.BatchPrpOpts),
    ("kite_job", kite.Job),
    ("kite_job_result", kite.JobResult),
    ("kite_run_opts", kite.RunOpts),
]


@pytest.fixture(scope="module")
def c_layouts(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "kite_batch.h"', "int main(void) {"]
    for cname, py in PAIRS:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = d / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "abi"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT}/include", str(src), "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.parametrize("cname,py", PAIRS, ids=[c for c, _ in PAIRS])
def test_ctypes_mirror_matches_c(c_layouts, cname, py):
    assert ctypes.sizeof(py) == c_layouts[(cname, "size")]
    for fname, _ in py._fields_:
        assert getattr(py, fname).offset == c_layouts[(cname, fname)], fname
