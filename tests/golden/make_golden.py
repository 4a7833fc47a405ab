#!/usr/bin/env python3
"""Writes the GMP golden fixtures under tests/golden/ (needs a host with GMP and a C compiler; no GPU).

  python tests/golden/make_golden.py plain_small     # 16 seeded n~65K tests: RES64 by GMP
  python tests/golden/make_golden.py chain           # x_i = 3^(k*2^i) mod N at chosen i
  python tests/golden/make_golden.py batch_subset N  # first N tests of the 256-test batch
  python tests/golden/make_golden.py batch_more      # the batch's other tests, smallest n first,
                                                     # each appended to _bin/batch_more.partial
  python tests/golden/make_golden.py batch_merge     # res64_batch_subset.csv += finished rows
  python tests/golden/make_golden.py single_large    # the seeded n~660K test
  python tests/golden/make_golden.py record_chain    # x_i of the record-size number (i = 50, 2000,
                                                     # 10000; GMP ~0.35 s per squaring at 25M bits),
                                                     # appended to chain_x.csv

GMP (mpz_powm, the semantics of the reference's tools/gmp_ref.c) runs through gmp_res64.c,
built here against the system libgmp. Outputs are CSV lines "k,n,base,RES64" (or
"k,n,base,i,low64,fnv64" for chain residues, fnv64 = FNV-1a-64 over the little-endian limbs,
(n+127)/64 limbs).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from ahnum_amd import workloads  # noqa: E402

BIN = os.path.join(HERE, "_bin", "gmp_res64")


def build():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    lib = "/usr/lib/x86_64-linux-gnu/libgmp.so.10"
    if not os.path.exists(lib):
        lib = "/opt/conda/lib/libgmp.so"
    subprocess.check_call(["gcc", "-O2", "-I/opt/conda/include", os.path.join(HERE, "gmp_res64.c"),
                           "-o", BIN, lib])


def gmp_lines(lines, jobs, partial=None):
    """Runs GMP on each input line; with `partial`, appends each result there as it completes."""
    def one(line):
        out = subprocess.run(["nice", "-n", "15", BIN], input=line + "\n", capture_output=True,
                             text=True, check=True).stdout.strip()
        if partial:
            with open(partial, "a") as f:
                f.write(out + "\n")
        return out
    with ThreadPoolExecutor(jobs) as ex:
        return list(ex.map(one, lines))


def write(name, header, rows):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        f.write(header + "\n")
        for r in rows:
            f.write(r + "\n")
    print("wrote", path, len(rows), "rows")


def read_rows(name):
    path = os.path.join(HERE, name)
    return [l.strip() for l in open(path) if l.strip() and not l.startswith("#")]


def main():
    what = sys.argv[1]
    jobs = int(os.environ.get("GOLDEN_JOBS", "4"))
    build()
    if what == "plain_small":
        tests = workloads.plain_small()
        rows = gmp_lines([f"{k} {n} 3" for k, n in tests], jobs)
        write("res64_plain_small.csv", "# k,n,base,RES64 (GMP mpz_powm): workloads.plain_small(16, seed=1)", rows)
    elif what == "chain":
        cases = []
        for k, n in workloads.plain_small()[:4]:
            for i in (0, 1, 7, 100, 1000):
                cases.append(f"{k} {n} 3 x {i}")
        for k, n in workloads.single_large() + workloads.batch_mixed()[:2]:
            for i in (0, 3, 50):
                cases.append(f"{k} {n} 3 x {i}")
        rows = gmp_lines(cases, jobs)
        write("chain_x.csv", "# k,n,base,i,low64,fnv64 of x_i = base^(k*2^i) mod N (GMP)", rows)
    elif what == "batch_subset":
        cnt = int(sys.argv[2])
        tests = workloads.batch_mixed()[:cnt]
        rows = gmp_lines([f"{k} {n} 3" for k, n in tests], jobs,
                         partial=os.path.join(HERE, "_bin", "batch_subset.partial"))
        write("res64_batch_subset.csv", f"# k,n,base,RES64 (GMP): first {cnt} of workloads.batch_mixed(256, seed=3)", rows)
    elif what == "batch_more":
        have = {tuple(r.split(",")[:2]) for r in read_rows("res64_batch_subset.csv")}
        todo = sorted((n, k) for k, n in workloads.batch_mixed() if (str(k), str(n)) not in have)
        gmp_lines([f"{k} {n} 3" for n, k in todo], jobs, partial=os.path.join(HERE, "_bin", "batch_more.partial"))
    elif what == "batch_merge":
        rows = {tuple(r.split(",")[:2]): r for r in read_rows("res64_batch_subset.csv")}
        part = os.path.join(HERE, "_bin", "batch_more.partial")
        if os.path.exists(part):
            for line in open(part):
                if line.strip():
                    rows[tuple(line.split(",")[:2])] = line.strip()
        order = [(str(k), str(n)) for k, n in workloads.batch_mixed()]
        out = [rows[key] for key in order if key in rows]
        write("res64_batch_subset.csv", f"# k,n,base,RES64 (GMP): {len(out)} of the 256 tests of "
              "workloads.batch_mixed(256, seed=3), in batch order", out)
    elif what == "record_chain":
        (k, n), = workloads.record()
        rows = gmp_lines([f"{k} {n} 3 x {i}" for i in (50, 2000, 10000)], jobs)
        have = read_rows("chain_x.csv")
        keep = [r for r in have if not r.startswith(f"{k},{n},")]
        path = os.path.join(HERE, "chain_x.csv")
        head = [l.rstrip("\n") for l in open(path) if l.startswith("#")]
        with open(path, "w") as f:
            f.write("\n".join(head + keep + rows) + "\n")
        print("wrote", path, len(keep + rows), "rows")
    elif what == "single_large":
        rows = gmp_lines([f"{k} {n} 3" for k, n in workloads.single_large()], 1,
                         partial=os.path.join(HERE, "_bin", "single_large.partial"))
        write("res64_single_large.csv", "# k,n,base,RES64 (GMP): workloads.single_large(seed=2)", rows)
    else:
        raise SystemExit("unknown fixture set " + what)


if __name__ == "__main__":
    main()
