"""Autotuner: measured cost of one squaring at every engine length.

Large engine: for each supported length N >= 2^16 the engine runs a short burst of
squarings of an exponent at 17 bits per digit and records the device time per squaring.
The learned length choice (nselect.choose_length) uses this table to run the fastest length
the round-off model deems safe -- not merely the shortest.

Batched engine (--batch): for each batched length N the GPU is filled with exponents of that
length (the largest prime within the batch cap of bits per digit, or the exactness bound for the NTT) and
the cost is the device time per exponent-squaring.  The library's length selector runs the
cheapest length within the bits-per-digit cap; the table is written as
csrc/ll_batch_costs.inc (the library's defaults) and batch_costs.json.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.autotune [--iters K]
writes length_costs.json next to this file;
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.autotune --batch
writes batch_costs.json and csrc/ll_batch_costs.inc (rebuild to make them the defaults).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

from . import llhip

HERE = os.path.dirname(os.path.abspath(__file__))
COSTS_PATH = os.path.join(HERE, "length_costs.json")


LOADED_JOBS = 4     # exponents in flight for the loaded cost (the large engine's streams)


def measure(iters: int = 600, min_n: int = 1 << 16) -> tuple[dict, dict]:
    """(alone, loaded): device microseconds per squaring at every large-engine length, with the
    exponent alone on the GPU and with LOADED_JOBS exponents of that length in one call."""
    alone, loaded = {}, {}
    for n in llhip.supported_lengths(llhip.FP64):
        if n < min_n:
            continue
        p = int(17 * n) | 1
        llhip.lucas_lehmer_n(p, 50, llhip.FP64, n, no_rollback=True)       # tables + warm-up
        r = llhip.lucas_lehmer_n(p, iters, llhip.FP64, n, no_rollback=True)
        alone[str(n)] = r["seconds"] / iters * 1e6
        llhip.batch_n([p] * LOADED_JOBS, iters, llhip.FP64, n, no_rollback=True)
        ms, _ = llhip.last_launch_stats()
        loaded[str(n)] = ms * 1e3 / (LOADED_JOBS * iters)
        print(f"N={n:8d} {alone[str(n)]:8.2f} us/squaring alone, {loaded[str(n)]:8.2f} with {LOADED_JOBS} in flight",
              flush=True)
    return alone, loaded


BATCH_JSON = os.path.join(HERE, "batch_costs.json")
BATCH_INC = os.path.join(HERE, "csrc", "ll_batch_costs.inc")


def _primes(lo: int, hi: int) -> list:
    if hi < 2:
        return []
    sv = bytearray([1]) * (hi + 1)
    sv[0:2] = b"\0\0"
    for i in range(2, int(hi ** 0.5) + 1):
        if sv[i]:
            sv[i * i::i] = bytearray(len(sv[i * i::i]))
    return [i for i in range(max(lo, 2), hi + 1) if sv[i]]


def batch_exponent(n: int, transform: int) -> int:
    """Largest prime the batched engine may run at length n (the batch cap for FP64)."""
    if transform == llhip.FP64:
        hi = int(llhip.get_batch_bpw() * n)
    else:
        hi = n * max(1, (63 - (n.bit_length() - 1)) // 2) - n // 2
    hi = min(hi, (1 << 32) // (n + 2) - 1)
    ps = _primes(max(65, hi - 4000), hi)
    return ps[-1]


def measure_batch(iters: int = 300, transforms=(llhip.FP64, llhip.NTT)) -> dict:
    out = {}
    for tr in transforms:
        lens = [n for n in llhip.supported_lengths(tr) if 32 <= n <= 8192 and llhip.get_batch_cost(n, tr) >= 0]
        for n in lens:
            if tr == llhip.NTT and n & (n - 1):
                continue
            m = n // 2 if tr == llhip.FP64 else n
            p = batch_exponent(n, tr)
            count = max(2048, min(65536, (8 << 20) // m))
            ps = [p] * count
            llhip.batch_n(ps[:64], 20, tr, n, no_rollback=True)              # tables + warm-up
            r = llhip.batch_n(ps, iters, tr, n, no_rollback=True)
            if not (r["n_used"] == n).all() or not (r["status"] == llhip.ST_OK).all():
                print(f"N={n}: not run at this length", flush=True)
                continue
            ms, _ = llhip.last_launch_stats()
            us = ms * 1e3 / (count * iters)
            out[(tr, n)] = us
            print(f"{'fp64' if tr == llhip.FP64 else 'ntt '} N={n:6d} p={p:7d} x{count:6d}: {us:9.5f} us/exponent-squaring",
                  flush=True)
    return out


def write_batch(costs: dict, iters: int) -> None:
    dev = llhip.device_info()
    with open(BATCH_JSON, "w") as f:
        json.dump({"device": dev["name"] or dev["arch"], "arch": dev["arch"], "iters": iters,
                   "us_per_exponent_squaring": {f"{'fp64' if tr == llhip.FP64 else 'ntt'}:{n}": v
                                                for (tr, n), v in sorted(costs.items())}}, f, indent=1)
    with open(BATCH_INC, "w") as f:
        f.write("// Batched-engine length costs, written by `python -m mersenneprime_python_cpp_gmp_benchmark_amd"
                ".autotune --batch`\n")
        f.write(f"// on {dev['name'] or dev['arch']} ({dev['arch']}): {{N, transform (0 FP64, 1 NTT), device "
                "microseconds per exponent-squaring at full occupancy}.\n")
        for (tr, n), v in sorted(costs.items()):
            f.write(f"    {{{n}, {tr}, {v:.6g}}},\n")


def load_costs(path: str = COSTS_PATH, loaded: bool = False) -> dict:
    """{N: us per squaring}; loaded: the cost with several exponents in flight (one call over
    many exponents) instead of one exponent alone."""
    with open(path) as f:
        d = json.load(f)
    key = "us_per_iter_loaded" if loaded and "us_per_iter_loaded" in d else "us_per_iter"
    return {int(k): float(v) for k, v in d[key].items()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=600)
    ap.add_argument("--out", default=COSTS_PATH)
    ap.add_argument("--batch", action="store_true", help="batched-engine lengths")
    a = ap.parse_args()
    if a.batch:
        it = a.iters if a.iters != 600 else 300
        write_batch(measure_batch(it), it)
        return 0
    alone, loaded = measure(a.iters)
    dev = llhip.device_info()
    with open(a.out, "w") as f:
        json.dump({"device": dev["name"] or dev["arch"], "arch": dev["arch"], "iters": a.iters,
                   "us_per_iter": alone, "loaded_jobs": LOADED_JOBS, "us_per_iter_loaded": loaded}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
