"""Round-off of the MI355X engine against bits per word, and the fit of the safe-bits model.

The reference measures, for a grid of FFT lengths L and multipliers k, the largest bits per word
whose maximum round-off over 300 squarings of a pseudo-random residue stays at or below 0.25
(tools/fit_bpw.py:19-45, with tools/sweep_err.c:23-36 for one measurement), and fits
bpw_max = A - B (log2 L - 12) - C log2 k to those limits (the model of src/ctx.c:26-33, used less a
0.2 margin). This script makes the same measurement on the GPU through the library (the kite
sweep measurement: x = 3, 64 squarings to spread the digits, round-off reset, `iters` squarings)
and prints:

  limit  L k bpw_max              largest bits per word (0.05 grid) with round-off <= threshold
  fit    A B C rms                least-squares fit of the limits to the model above
  table  L k bpw maxerr           round-off at the selector's own bound kite_safe_bpw(L, k)

Usage: python -m ahnum_amd.fit_bpw [--threshold 0.25] [--iters 300] [--lengths L,L,...] [--ks k,k,...]
                                   [--no-limits] [--no-table]
"""
import argparse
import math
import sys

import numpy as np

from ahnum_amd import kite

DEFAULT_LENGTHS = [4096, 16384, 49152, 65536, 262144, 1048576]
DEFAULT_KS = [1, 3, 31, 255, 2047, 9999]


def maxerr(L, k, bpw, iters):
    """Max round-off of `iters` squarings at length L for N = k*2^n-1, n = bpw*L; None if the
    library refuses that length for that number."""
    n = int(bpw * L)
    try:
        c = kite.Ctx(k, n, fft_len=L)
    except RuntimeError:
        return None
    try:
        x = c.num(3)
        x.square(64)
        c.reset_maxerr()
        x.square(iters)
        e = c.maxerr()
        x.free()
        return e
    finally:
        c.close()


def limit(L, k, thr, iters):
    """Largest bits per word (0.05 grid) whose round-off stays <= thr, by bisection from the
    model's guess (tools/fit_bpw.py:29-38)."""
    guess = (53 - 6 - 0.5 * math.log2(L) - math.log2(k)) / 2
    lo, hi = guess - 3, guess + 4
    while hi - lo > 0.05:
        mid = (lo + hi) / 2
        e = maxerr(L, k, mid, iters)
        if e is not None and e <= thr:
            lo = mid
        else:
            hi = mid
    return lo


def fit(rows):
    """Least squares of bpw_max = A - B (log2 L - 12) - C log2 k over (L, k, bpw_max) rows."""
    X = np.array([[1.0, -(math.log2(L) - 12), -math.log2(k) if k > 1 else 0.0] for L, k, _ in rows])
    y = np.array([b for _, _, b in rows])
    coef, *_ = np.linalg.lstsq(X, y, rcond=None)
    rms = float(np.sqrt(np.mean((X @ coef - y) ** 2)))
    return coef, rms


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--threshold", type=float, default=0.25)
    ap.add_argument("--iters", type=int, default=300)
    ap.add_argument("--lengths", default=",".join(map(str, DEFAULT_LENGTHS)))
    ap.add_argument("--ks", default=",".join(map(str, DEFAULT_KS)))
    ap.add_argument("--no-limits", action="store_true")
    ap.add_argument("--no-table", action="store_true")
    a = ap.parse_args(argv)
    Ls = [int(v) for v in a.lengths.split(",")]
    ks = [int(v) for v in a.ks.split(",")]
    lib = kite.lib()
    rows = []
    if not a.no_limits:
        for L in Ls:
            for k in ks:
                b = limit(L, k, a.threshold, a.iters)
                rows.append((L, k, b))
                print(f"limit L={L} k={k} bpw_max={b:.2f} model={lib.kite_safe_bpw(L, k) + 0.2:.2f}", flush=True)
        (A, B, C), rms = fit(rows)
        print(f"fit A={A:.3f} B={B:.3f} C={C:.3f} rms={rms:.3f} (reference model: A=21.44 B=0.305 C=0.5)",
              flush=True)
    if not a.no_table:
        for L in Ls:
            for k in ks:
                bpw = lib.kite_safe_bpw(L, k)
                e = maxerr(L, k, bpw, a.iters)
                print(f"table L={L} k={k} bpw={bpw:.3f} maxerr={'refused' if e is None else '%.4f' % e}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
