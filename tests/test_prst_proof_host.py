"""Host parts of kite-prst's PRST proof module (ahnum_amd/csrc/prst_proof.cpp): MD5 against the
RFC 1321 test suite, PRST's fingerprints against the file names the reference's
tests/prst_compat.sh expects, point positions, the state-file layout of docs/prst-compat.md and
the challenge's form. A small C++ driver is compiled against the module; nothing here needs a GPU."""
import hashlib
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
#include <cstring>
#include <string>
#include "prst_proof.h"
int main(int argc, char **argv) {
    const std::string op = argv[1];
    if (op == "md5") {  // md5 <text>
        printf("%s\n", prst::Md5::hex(argv[2], strlen(argv[2])).c_str());
    } else if (op == "md5rep") {  // md5rep <text> <times>, streamed in uneven pieces
        prst::Md5 m;
        const int times = atoi(argv[3]);
        for (int i = 0; i < times; i++) m.update(argv[2], strlen(argv[2]));
        unsigned char d[16];
        m.final(d);
        for (int i = 0; i < 16; i++) printf("%02x", d[i]);
        printf("\n");
    } else if (op == "fp") {  // fp k n a
        const uint64_t k = strtoull(argv[2], 0, 10), n = strtoull(argv[3], 0, 10);
        printf("%u %u\n", prst::fingerprint(k, n), prst::proof_fingerprint(k, n, (uint32_t)atoi(argv[4])));
    } else if (op == "points") {  // points iterations count
        std::vector<uint32_t> pos;
        const int M = prst::points((uint32_t)strtoul(argv[2], 0, 10), atoi(argv[3]), &pos);
        printf("%d", M);
        for (uint32_t p : pos) printf(" %u", p);
        printf("\n");
    } else if (op == "write") {  // write path type fp iteration words...
        prst::Giant g;
        for (int i = 6; i < argc; i++) g.w.push_back((uint32_t)strtoul(argv[i], 0, 0));
        return prst::write_state(argv[2], atoi(argv[3]), (uint32_t)strtoul(argv[4], 0, 10), atoi(argv[5]), {g, prst::Giant()});
    } else if (op == "read") {  // read path type fp
        std::vector<prst::Giant> g;
        int32_t it;
        if (prst::read_state(argv[2], atoi(argv[3]), (uint32_t)strtoul(argv[4], 0, 10), &it, &g, 3)) return 3;
        printf("%d %zu %zu %zu", it, g[0].w.size(), g[1].w.size(), g[2].w.size());
        for (uint32_t w : g[0].w) printf(" %u", w);
        printf("\n");
    } else if (op == "challenge") {  // challenge fp y0 d0
        prst::Giant Y, D;
        Y.w.push_back((uint32_t)strtoul(argv[3], 0, 10));
        D.w.push_back((uint32_t)strtoul(argv[4], 0, 10));
        const unsigned __int128 h = prst::challenge((uint32_t)strtoul(argv[2], 0, 10), Y, D);
        printf("%llu %llu\n", (unsigned long long)(h >> 64), (unsigned long long)h);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def drv(tmp_path_factory):
    d = tmp_path_factory.mktemp("prst")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT}/include", f"-I{CSRC}", str(src),
                    os.path.join(CSRC, "prst_proof.cpp"), "-o", str(exe), f"-L{LIB}", "-lkite",
                    f"-Wl,-rpath,{LIB}"], check=True, timeout=300)

    def run(*args):
        r = subprocess.run([str(exe)] + [str(a) for a in args], capture_output=True, text=True, timeout=60)
        return r.returncode, r.stdout.strip()
    return run


@pytest.mark.parametrize("text,digest", [  # RFC 1321, appendix A.5
    ("", "d41d8cd98f00b204e9800998ecf8427e"),
    ("a", "0cc175b9c0f1b6a831c399e269772661"),
    ("abc", "900150983cd24fb0d6963f7d28e17f72"),
    ("message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
    ("abcdefghijklmnopqrstuvwxyz", "c3fcd3d76192e4007dfb496cca67e13b"),
    ("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789", "d174ab98d277d9f5a5611c2c9f419d9f"),
    ("12345678901234567890123456789012345678901234567890123456789012345678901234567890",
     "57edf4a22be3c955ac49da2e2107b67a"),
])
def test_md5_rfc1321(drv, text, digest):
    assert drv("md5", text) == (0, digest)


def test_md5_streamed_across_blocks(drv):
    for piece, times in [("x" * 63, 5), ("0123456789", 100), ("abc", 1000)]:
        assert drv("md5rep", piece, times)[1] == hashlib.md5((piece * times).encode()).hexdigest()


def test_fingerprints_name_the_reference_files(drv):
    """tests/prst_compat.sh: 1203*2^90001-1's files are prst_1107603830.*; fingerprint = N mod
    3417905339; the proof fingerprint is the first word of MD5(fp || "3.90001")."""
    rc, out = drv("fp", 1203, 90001, 3)
    fp, pfp = map(int, out.split())
    assert fp == 1107603830 == (1203 * 2 ** 90001 - 1) % 3417905339
    assert pfp == struct.unpack("<I", hashlib.md5(struct.pack("<I", fp) + b"3.90001").digest()[:4])[0]


@pytest.mark.parametrize("iters,count", [(90001, 512), (1000, 8), (20000, 16), (659321, 64), (1 << 20, 4)])
def test_points_subdivide_the_iterations(drv, iters, count):
    """Point positions: 0 and `iters` at the ends, increasing, each midpoint splitting its range into
    halves within a few iterations of each other, recursively (PRST's subdivision rounds each level
    by the iteration count's bits, so the halves of an inner range can differ by more than one; the
    exact positions are pinned by the product files' MD5s in the GPU proof test)."""
    vals = list(map(int, drv("points", iters, count)[1].split()))
    pos = vals[1:]
    assert len(pos) == count + 1 and pos[0] == 0 and pos[-1] == iters
    assert all(b > a for a, b in zip(pos, pos[1:]))
    def check(lo, hi):
        if hi - lo < 2:
            return
        mid = (lo + hi) // 2
        assert abs((pos[mid] - pos[lo]) - (pos[hi] - pos[mid])) <= count.bit_length()
        check(lo, mid)
        check(mid, hi)
    check(0, count)


def test_state_file_layout_and_sidecar(drv, tmp_path):
    """docs/prst-compat.md: u32 magic 0x9f2b3cd4, u32 appid 4 | type << 16, u32 fingerprint,
    i32 iteration, then giants as i32 word count + words; the .md5 sidecar is the file's hex MD5."""
    p = tmp_path / "prst_1.prod.3"
    assert drv("write", p, 3, 77, 3, "0x11223344", "5")[0] == 0
    raw = p.read_bytes()
    assert raw == struct.pack("<IIIiI", 0x9F2B3CD4, 4 | 3 << 16, 77, 3, 2) + struct.pack("<II", 0x11223344, 5) + struct.pack("<i", 0)
    assert (tmp_path / "prst_1.prod.3.md5").read_text() == hashlib.md5(raw).hexdigest()
    assert drv("read", p, 3, 77) == (0, f"3 2 0 0 {0x11223344} 5")
    assert drv("read", p, 3, 78)[0] != 0          # wrong fingerprint
    assert drv("read", p, 1, 77)[0] != 0          # wrong type
    (tmp_path / "prst_1.prod.3.md5").write_text("0" * 32)
    assert drv("read", p, 3, 77)[0] != 0          # sidecar mismatch


def test_challenge_form(drv):
    """MD5(fp || Y || D), its first two words as a 64-bit number, made odd and free of prime factors
    below 1000, stepping by 2."""
    primes = [p for p in range(3, 1000) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    for fp, y, d in [(1107603830, 5, 7), (1, 2, 3), (3417905338, 123456789, 987654321)]:
        hi, lo = map(int, drv("challenge", fp, y, d)[1].split())
        h = hi << 64 | lo
        dig = hashlib.md5(struct.pack("<III", fp, y, d)).digest()
        w0, w1 = struct.unpack("<II", dig[:8])
        start = (w1 or 1) << 32 | (w0 | 1)
        assert h >= start and (h - start) % 2 == 0 and h % 2 == 1
        assert all(h % p for p in primes)
        assert all(any(c % p == 0 for p in primes) for c in range(start, h, 2))
