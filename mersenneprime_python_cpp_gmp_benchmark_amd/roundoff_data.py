"""Measures the training data of the round-off model on the GPU.

For every training exponent of tests/golden/c4_exponents.csv and every supported FP64 length
with 14.5..21.5 bits per digit, runs the first 2000 squarings (no rollback) and records the
worst round-off.  Output: roundoff_train.csv (p, N, maxerr, status).

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.roundoff_data [--limit K]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from . import llhip
from .nselect import DATA_PATH, WINDOW_ITERS, candidate_lengths

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def exponents(split: str):
    rows = [l.strip().split(",") for l in open(os.path.join(ROOT, "tests", "golden", "c4_exponents.csv"))][1:]
    return [int(p) for p, s in rows if s == split]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--limit", type=int, default=0)
    ap.add_argument("--out", default=DATA_PATH)
    ap.add_argument("--lo", type=float, default=14.5)
    ap.add_argument("--hi", type=float, default=21.5)
    a = ap.parse_args()
    ps = exponents("train")
    if a.limit:
        ps = ps[:: max(1, len(ps) // a.limit)][: a.limit]
    cands = candidate_lengths()
    t0 = time.time()
    n_runs = 0
    with open(a.out, "w") as f:
        f.write("p,n,maxerr,status\n")
        for i, p in enumerate(ps):
            for n in cands:
                if not (a.lo <= p / n <= a.hi):
                    continue
                r = llhip.lucas_lehmer_n(p, WINDOW_ITERS, llhip.FP64, n, no_rollback=True)
                f.write(f"{p},{n},{r['maxerr']:.6g},{r['status']}\n")
                n_runs += 1
            f.flush()
            print(f"{i + 1}/{len(ps)} p={p} runs={n_runs} {time.time() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
