"""Autotuner: measured cost of one squaring at every large-engine FP64 length.

For each supported length N >= 2^16 the engine runs a short burst of squarings of an
exponent at 17 bits per digit and records the device time per squaring.  The learned
length choice (nselect.choose_length) uses this table to run the fastest length the
round-off model deems safe -- not merely the shortest.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.autotune [--iters K]
writes length_costs.json next to this file.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

from . import llhip

HERE = os.path.dirname(os.path.abspath(__file__))
COSTS_PATH = os.path.join(HERE, "length_costs.json")


def measure(iters: int = 600, min_n: int = 1 << 16) -> dict:
    out = {}
    for n in llhip.supported_lengths(llhip.FP64):
        if n < min_n:
            continue
        p = int(17 * n) | 1
        llhip.lucas_lehmer_n(p, 50, llhip.FP64, n, no_rollback=True)       # tables + warm-up
        r = llhip.lucas_lehmer_n(p, iters, llhip.FP64, n, no_rollback=True)
        out[str(n)] = r["seconds"] / iters * 1e6
        print(f"N={n:8d} {out[str(n)]:8.2f} us/squaring", flush=True)
    return out


def load_costs(path: str = COSTS_PATH) -> dict:
    with open(path) as f:
        return {int(k): float(v) for k, v in json.load(f)["us_per_iter"].items()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=600)
    ap.add_argument("--out", default=COSTS_PATH)
    a = ap.parse_args()
    costs = measure(a.iters)
    dev = llhip.device_info()
    with open(a.out, "w") as f:
        json.dump({"device": dev["name"] or dev["arch"], "arch": dev["arch"], "iters": a.iters,
                   "us_per_iter": costs}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
