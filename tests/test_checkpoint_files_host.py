"""Checkpoint files written and read by the library (ahnum_amd/csrc/prp.cpp, llr.cpp), byte for byte
against the reference's layouts: KITECKP1 = {magic[8], k, n, base, iter, nlimbs, fft_len} (u64
each), x limbs, d limbs, FNV-1a-64 of everything before it (src/prp.c:146-212); the old magic
"AHNCKPT1" is read too (src/prp.c:163-170). KITELLR1 = {magic[8], k, n, P, iter, nlimbs, fft_len,
run, prev[2]}, x limbs, FNV-1a-64 (src/llr.c:92-146). No GPU: a small C++ driver calls the library's
file functions, and the public kite_*_checkpoint_iter readers are called through ctypes."""
import ctypes
import os
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ahnum_amd", "csrc")
LIB = os.path.join(ROOT, "ahnum_amd", "lib")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(LIB, "libkite.so")), reason="libkite.so not built")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "kite_internal.h"
static uint64_t u(const char *s) { return strtoull(s, 0, 0); }
int main(int argc, char **argv) {
    const std::string op = argv[1];
    const char *path = argv[2];
    if (op == "pwrite" || op == "pread") {  // path k n base nl [iter L]
        const size_t nl = u(argv[6]);
        std::vector<uint64_t> x(nl), d(nl);
        if (op == "pwrite") {
            for (size_t i = 0; i < nl; i++) {
                x[i] = 0x0101010101010101ull * (i + 1);
                d[i] = ~x[i];
            }
            return kite::prp_checkpoint_write(path, u(argv[3]), u(argv[4]), (uint32_t)u(argv[5]), u(argv[7]),
                                              (uint32_t)u(argv[8]), x.data(), d.data(), nl) ? 1 : 0;
        }
        uint64_t it = 0;
        uint32_t L = 0;
        if (kite::prp_checkpoint_read(path, u(argv[3]), u(argv[4]), (uint32_t)u(argv[5]), &it, &L, x.data(), d.data(), nl))
            return 1;
        printf("%llu %u %llu %llu\n", (unsigned long long)it, L, (unsigned long long)x[0], (unsigned long long)d[nl - 1]);
        return 0;
    }
    // lwrite path k n nl P iter L run prev0 prev1 | lread path k n nl
    const size_t nl = u(argv[5]);
    std::vector<uint64_t> x(nl);
    kite::LlrCheckpoint c;
    if (op == "lwrite") {
        for (size_t i = 0; i < nl; i++) x[i] = 0x1111111111111111ull * (i + 3);
        c.P = (uint32_t)u(argv[6]);
        c.iter = u(argv[7]);
        c.L = (uint32_t)u(argv[8]);
        c.run = (int)u(argv[9]);
        c.prev[0] = u(argv[10]);
        c.prev[1] = u(argv[11]);
        return kite::llr_checkpoint_write(path, u(argv[3]), u(argv[4]), c, x.data(), nl) ? 1 : 0;
    }
    if (kite::llr_checkpoint_read(path, u(argv[3]), u(argv[4]), &c, x.data(), nl)) return 1;
    printf("%llu %u %u %d %llu %llu %llu\n", (unsigned long long)c.iter, c.P, c.L, c.run, (unsigned long long)c.prev[0],
           (unsigned long long)c.prev[1], (unsigned long long)x[0]);
    return 0;
}
"""


def fnv1a64(data, h=0xCBF29CE484222325):
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv1a64_known_answers():
    # the published FNV-1a 64-bit test vectors (Fowler/Noll/Vo): "", "a", "foobar"
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8


@pytest.fixture(scope="module")
def drv(tmp_path_factory):
    d = tmp_path_factory.mktemp("ckpt")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT}/include", f"-I{CSRC}", str(src), "-o", str(exe),
                    f"-L{LIB}", "-lkite", f"-Wl,-rpath,{LIB}"], check=True, timeout=300)

    def run(*args):
        r = subprocess.run([str(exe)] + [str(a) for a in args], capture_output=True, text=True, timeout=60)
        return r.returncode, r.stdout.strip()
    return run


def limbs(vals):
    return b"".join(struct.pack("<Q", v & 0xFFFFFFFFFFFFFFFF) for v in vals)


def test_fermat_file_layout_and_checks(drv, tmp_path):
    k, n, base, it, L, nl = 9357, 20000, 3, 8000, 2048, (20000 + 127) // 64
    p = tmp_path / "kite_1.3.ckpt"
    assert drv("pwrite", p, k, n, base, nl, it, L)[0] == 0
    raw = p.read_bytes()
    x = [0x0101010101010101 * (i + 1) for i in range(nl)]
    body = struct.pack("<8s6Q", b"KITECKP1", k, n, base, it, nl, L) + limbs(x) + limbs([~v for v in x])
    assert raw == body + struct.pack("<Q", fnv1a64(body))
    assert not (tmp_path / "kite_1.3.ckpt.tmp").exists()  # written through a renamed temporary
    assert drv("pread", p, k, n, base, nl) == (0, f"{it} {L} {x[0]} {(~x[-1]) & 0xFFFFFFFFFFFFFFFF}")
    lib = ctypes.CDLL(os.path.join(LIB, "libkite.so"))
    got = ctypes.c_uint64()
    assert lib.kite_prp_checkpoint_iter(str(p).encode(), ctypes.c_uint64(k), ctypes.c_uint64(n), ctypes.c_uint32(base),
                                        ctypes.byref(got)) == 0 and got.value == it
    # refused: another number, base or limb count; a flipped byte
    assert drv("pread", p, k + 2, n, base, nl)[0] != 0
    assert drv("pread", p, k, n, 5, nl)[0] != 0
    assert drv("pread", p, k, n, base, nl + 1)[0] != 0
    bad = bytearray(raw)
    bad[100] ^= 1
    p.write_bytes(bytes(bad))
    assert drv("pread", p, k, n, base, nl)[0] != 0
    # the old magic, with its checksum, reads as the same state
    old = b"AHNCKPT1" + body[8:]
    p.write_bytes(old + struct.pack("<Q", fnv1a64(old)))
    assert drv("pread", p, k, n, base, nl)[0] == 0


def test_llr_file_layout_and_checks(drv, tmp_path):
    k, n, nl, P, it, L, run, p0, p1 = 33, 99999, (99999 + 127) // 64, 5, 41000, 6144, 1, 0x3FAF838186E8A836, 0
    p = tmp_path / "kite_1.llr.ckpt"
    assert drv("lwrite", p, k, n, nl, P, it, L, run, p0, p1)[0] == 0
    raw = p.read_bytes()
    x = [0x1111111111111111 * (i + 3) for i in range(nl)]
    body = struct.pack("<8s9Q", b"KITELLR1", k, n, P, it, nl, L, run, p0, p1) + limbs(x)
    assert raw == body + struct.pack("<Q", fnv1a64(body))
    assert drv("lread", p, k, n, nl) == (0, f"{it} {P} {L} {run} {p0} {p1} {x[0] & 0xFFFFFFFFFFFFFFFF}")
    lib = ctypes.CDLL(os.path.join(LIB, "libkite.so"))
    got, r = ctypes.c_uint64(), ctypes.c_int()
    assert lib.kite_llr_checkpoint_iter(str(p).encode(), ctypes.c_uint64(k), ctypes.c_uint64(n), ctypes.byref(got),
                                        ctypes.byref(r)) == 0 and (got.value, r.value) == (it, run)
    assert drv("lread", p, k + 2, n, nl)[0] != 0
    bad = bytearray(raw)
    bad[-9] ^= 0x80
    p.write_bytes(bytes(bad))
    assert drv("lread", p, k, n, nl)[0] != 0
