"""Autotuner: measured cost of one squaring at every engine length, and the large engine's
schedule (H x W split and column-pass block width) per length.

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

Large-engine schedules (--schedules): every schedule of every large length (both transforms)
runs a burst of squarings alone on the GPU; the fastest becomes the length's schedule
(csrc/ll_large_sched_defaults.inc, the library's default after a rebuild; large_schedules.json
keeps every measurement), and the length costs are then measured at the chosen schedules.

Every cost table records a hash of the kernel sources it was measured on (comments and blank
space ignored); kernel_source_hash() recomputes it, so a test can tell a stale table.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import sys

from . import llhip

HERE = os.path.dirname(os.path.abspath(__file__))
COSTS_PATH = os.path.join(HERE, "length_costs.json")
CSRC = os.path.join(HERE, "csrc")
SCHED_INC = os.path.join(CSRC, "ll_large_sched_defaults.inc")
SCHED_JSON = os.path.join(HERE, "large_schedules.json")

# the sources whose code decides each table's costs
KERNEL_SOURCES = {
    "large": ["ll_large_impl.hpp", "ll_large_scheds.inc", "ll_large_sched_defaults.inc", "ll_dft.hpp",
              "ll_common.hpp"],
    "batch": ["ll_batch.hip", "ll_batch_sched_defaults.inc", "ll_dft.hpp", "ll_common.hpp"],
}


def kernel_source_hash(kind: str) -> str:
    """sha256 (16 hex digits) of the kernel sources of one table, comments and blank space removed."""
    h = hashlib.sha256()
    for name in KERNEL_SOURCES[kind]:
        with open(os.path.join(CSRC, name)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        h.update(name.encode())
        h.update(re.sub(r"\s+", "", text).encode())
    return h.hexdigest()[:16]


def prime_at_most(hi: int) -> int:
    """Largest prime <= hi (deterministic Miller-Rabin, exact below 3.3e24)."""
    def is_prime(n: int) -> bool:
        if n < 2:
            return False
        for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            if n % q == 0:
                return n == q
        d, r = n - 1, 0
        while d % 2 == 0:
            d //= 2
            r += 1
        for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            x = pow(a, d, n)
            if x in (1, n - 1):
                continue
            for _ in range(r - 1):
                x = x * x % n
                if x == n - 1:
                    break
            else:
                return False
        return True
    while not is_prime(hi):
        hi -= 1
    return hi


LOADED_JOBS = 4     # exponents in flight for the loaded cost (the large engine's streams)


def per_squaring_us(run, iters: int, reps: int = 2) -> float:
    """Device microseconds per squaring of `run(k)` (device seconds of k squarings): the slope
    between bursts of iters // 4 and iters squarings, so what a run costs once (set-up, graph
    capture, normalisation) cancels; best of `reps`."""
    k0 = max(iters // 4, 1)
    t0 = min(run(k0) for _ in range(reps))
    t1 = min(run(iters) for _ in range(reps))
    return (t1 - t0) / (iters - k0) * 1e6


def measure(iters: int = 800, min_n: int = 1 << 16) -> tuple[dict, dict]:
    """(alone, loaded): device microseconds per squaring at every large-engine length, with the
    exponent alone on the GPU and with LOADED_JOBS exponents of that length in one call."""
    alone, loaded = {}, {}

    def one(p, n):
        return lambda k: llhip.lucas_lehmer_n(p, k, llhip.FP64, n, no_rollback=True)["seconds"]

    def four(p, n):
        def run(k):
            llhip.batch_n([p] * LOADED_JOBS, k, llhip.FP64, n, no_rollback=True)
            return llhip.last_launch_stats()[0] / 1e3
        return run

    for n in llhip.supported_lengths(llhip.FP64):
        if n < min_n:
            continue
        p = prime_at_most(17 * n)
        llhip.lucas_lehmer_n(p, 50, llhip.FP64, n, no_rollback=True)       # tables + warm-up
        alone[str(n)] = per_squaring_us(one(p, n), iters)
        loaded[str(n)] = per_squaring_us(four(p, n), iters) / LOADED_JOBS
        print(f"N={n:8d} {alone[str(n)]:8.2f} us/squaring alone, {loaded[str(n)]:8.2f} with {LOADED_JOBS} in flight",
              flush=True)
    return alone, loaded


def measure_schedules(iters: int = 800, transforms=(llhip.FP64, llhip.NTT), reps: int = 2) -> dict:
    """{(transform, N): [(W, H, CW, us per squaring)]}: every schedule of every large length, the
    exponent (the largest prime <= 16.5 N) alone on the GPU, best of `reps` bursts."""
    out = {}
    for tr in transforms:
        for n in llhip.supported_lengths(tr):
            sch = llhip.large_schedules(n, tr)
            if not sch:
                continue
            p = prime_at_most(int(16.5 * n))
            rows = []
            try:
                for s, (w, h, cw) in enumerate(sch):
                    llhip.set_large_schedule(n, s, tr)
                    llhip.lucas_lehmer_n(p, 50, tr, n, no_rollback=True)       # tables, graphs, warm-up
                    us = per_squaring_us(lambda k: llhip.lucas_lehmer_n(p, k, tr, n, no_rollback=True)["seconds"],
                                         iters, reps)
                    rows.append((w, h, cw, us))
            finally:
                llhip.set_large_schedule(n, -1, tr)
            out[(tr, n)] = rows
            best = min(rows, key=lambda r: r[3])
            print(f"{'fp64' if tr == llhip.FP64 else 'ntt '} N={n:8d}: " +
                  "  ".join(f"{w}x{h}/{cw} {us:6.2f}" for w, h, cw, us in rows) +
                  f"  -> {best[0]}x{best[1]}/{best[2]}", flush=True)
    return out


def write_schedules(meas: dict, iters: int) -> dict:
    """Writes the fastest schedule per length (SCHED_INC, SCHED_JSON); returns {(tr, N): index}."""
    dev = llhip.device_info()
    choice = {}
    with open(SCHED_INC, "w") as f:
        f.write("// Measured schedule of each large-engine length: LLD(transform, N, W, H, CW), transform 0 FP64,\n"
                "// 1 NTT.  Written by `python -m mersenneprime_python_cpp_gmp_benchmark_amd.autotune --schedules`\n"
                f"// on {dev['name'] or dev['arch']} ({dev['arch']}): the fastest schedule of ll_large_scheds.inc with\n"
                f"// one exponent alone on the GPU ({iters} squarings); a length without a row runs the first one.\n")
        for (tr, n), rows in sorted(meas.items()):
            i = min(range(len(rows)), key=lambda k: rows[k][3])
            choice[(tr, n)] = i
            w, h, cw, us = rows[i]
            f.write(f"LLD({tr}, {n}, {w}, {h}, {cw})\n")
    with open(SCHED_JSON, "w") as f:
        json.dump({"device": dev["name"] or dev["arch"], "arch": dev["arch"], "iters": iters,
                   "kernel_hash": kernel_source_hash("large"),
                   "us_per_squaring": {f"{'fp64' if tr == llhip.FP64 else 'ntt'}:{n}":
                                       {f"{w}x{h}/{cw}": round(us, 3) for w, h, cw, us in rows}
                                       for (tr, n), rows in sorted(meas.items())}}, f, indent=1)
    return choice


BATCH_JSON = os.path.join(HERE, "batch_costs.json")
BATCH_INC = os.path.join(HERE, "csrc", "ll_batch_costs.inc")
BATCH_SCHED_INC = os.path.join(HERE, "csrc", "ll_batch_sched_defaults.inc")


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


def measure_batch(iters: int = 300, transforms=(llhip.FP64, llhip.NTT)) -> tuple[dict, dict]:
    """(costs, schedules): costs[(tr, n)] = device microseconds per exponent-squaring with the GPU
    full of exponents of length n at its fastest schedule; schedules[(tr, n)] = [(radices, us)]
    for every schedule of a length that has several (the choice is the fastest)."""
    out, scheds = {}, {}
    for tr in transforms:
        lens = [n for n in llhip.supported_lengths(tr) if 32 <= n <= 8192 and llhip.get_batch_cost(n, tr) >= 0]
        for n in lens:
            if tr == llhip.NTT and n & (n - 1):
                continue
            m = n // 2 if tr == llhip.FP64 else n
            p = batch_exponent(n, tr)
            count = max(2048, min(65536, (8 << 20) // m))
            ps = [p] * count
            radices = llhip.batch_schedules(n, tr)
            rows = []
            try:
                for s in range(len(radices)):
                    llhip.set_batch_schedule(n, s, tr)
                    llhip.batch_n(ps[:64], 20, tr, n, no_rollback=True)              # tables + warm-up
                    r = llhip.batch_n(ps, iters, tr, n, no_rollback=True)
                    if not (r["n_used"] == n).all() or not (r["status"] == llhip.ST_OK).all():
                        rows = []
                        break
                    ms, _ = llhip.last_launch_stats()
                    rows.append((radices[s], ms * 1e3 / (count * iters)))
            finally:
                llhip.set_batch_schedule(n, -1, tr)
            if not rows:
                print(f"N={n}: not run at this length", flush=True)
                continue
            best = min(range(len(rows)), key=lambda i: rows[i][1])
            out[(tr, n)] = rows[best][1]
            if len(rows) > 1:
                scheds[(tr, n)] = (best, rows)
            print(f"{'fp64' if tr == llhip.FP64 else 'ntt '} N={n:6d} p={p:7d} x{count:6d}: " +
                  "  ".join(f"{'x'.join(map(str, rad))} {us:9.5f}" for rad, us in rows) + " us/exponent-squaring",
                  flush=True)
    return out, scheds


def write_batch(costs: dict, scheds: dict, iters: int) -> None:
    dev = llhip.device_info()
    with open(BATCH_SCHED_INC, "w") as f:
        f.write("// Measured schedule of each batched length: LLB(transform, N, index into llh_batch_schedules),\n"
                "// transform 0 FP64, 1 NTT.  Written by `python -m mersenneprime_python_cpp_gmp_benchmark_amd"
                ".autotune --batch`\n"
                f"// on {dev['name'] or dev['arch']} ({dev['arch']}): the fastest radix order with the GPU full of "
                "exponents of that length.\n")
        for (tr, n), (best, _) in sorted(scheds.items()):
            f.write(f"LLB({tr}, {n}, {best})\n")
    with open(BATCH_JSON, "w") as f:
        json.dump({"device": dev["name"] or dev["arch"], "arch": dev["arch"], "iters": iters,
                   "kernel_hash": kernel_source_hash("batch"),
                   "us_per_exponent_squaring": {f"{'fp64' if tr == llhip.FP64 else 'ntt'}:{n}": v
                                                for (tr, n), v in sorted(costs.items())},
                   "schedules": {f"{'fp64' if tr == llhip.FP64 else 'ntt'}:{n}":
                                 {"x".join(map(str, rad)): us for rad, us in rows}
                                 for (tr, n), (_, rows) in sorted(scheds.items())}}, f, indent=1)
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
    ap.add_argument("--iters", type=int, default=800)
    ap.add_argument("--out", default=COSTS_PATH)
    ap.add_argument("--batch", action="store_true", help="batched-engine lengths")
    ap.add_argument("--schedules", action="store_true",
                    help="large engine: pick each length's schedule, then measure the length costs at it")
    a = ap.parse_args()
    if a.batch:
        it = a.iters if a.iters != 800 else 300
        costs, scheds = measure_batch(it)
        write_batch(costs, scheds, it)
        return 0
    if a.schedules:
        choice = write_schedules(measure_schedules(), 800)
        for (tr, n), i in choice.items():           # the costs below run at the chosen schedules
            llhip.set_large_schedule(n, i, tr)
    alone, loaded = measure(a.iters)
    dev = llhip.device_info()
    with open(a.out, "w") as f:
        json.dump({"device": dev["name"] or dev["arch"], "arch": dev["arch"], "iters": a.iters,
                   "kernel_hash": kernel_source_hash("large"),
                   "us_per_iter": alone, "loaded_jobs": LOADED_JOBS, "us_per_iter_loaded": loaded}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
