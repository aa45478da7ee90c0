"""Range driver for the device engines.

The reference drives its CPU engines from a pool of Python threads, one exponent per thread, and
sends its GPU exponents through `lucaslehmer_cuda.lucas_lehmer_batch` in chunks that share a
transform length (PerfNumMultiCLL.py:136-164).  Here every engine runs on the GPU and takes a
whole list per call, so the driver is a loop over ascending chunks of the exponent list: one
device call per chunk, its rows printed as the chunk returns.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -r 100001          # primes p <= 100001
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -p 11213            # one exponent
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -l 521 607 1279     # a list
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -n 2 10001          # count primes only
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -r 10001 --json a.json
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -r 10001 --verify a.json --loop 3

Engines: `fft` (default) is the FP64 IBDWT with its per-squaring round-off check; `gmp` and `cuda`
are the exact Goldilocks NTT (`cuda` keeps the reference's default of 2048 exponents per call).
An exponent a call hands back with seconds < 0 is redone alone through `lucaslehmer`, whose FP64
functions redo rounding trouble on the exact engine.  Kept from the reference as its interface:
the option names, the table row, the summary line, the JSON keys, and the exit codes (0 ok, 1 bad
arguments, 2 a known Mersenne exponent in range not found, 3 a Res64 mismatch under --verify).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Callable, Iterable

MAX_WORKERS = 256

# Exponents of the 52 known Mersenne primes (the public GIMPS list).
KNOWN = (2, 3, 5, 7, 13, 17, 19, 31, 61, 89, 107, 127, 521, 607, 1279, 2203, 2281, 3217, 4253, 4423,
         9689, 9941, 11213, 19937, 21701, 23209, 44497, 86243, 110503, 132049, 216091, 756839, 859433,
         1257787, 1398269, 2976221, 3021377, 6972593, 13466917, 20996011, 24036583, 25964951, 30402457,
         32582657, 37156667, 42643801, 43112609, 57885161, 74207281, 77232917, 82589933, 136279841)

Engine = Callable[[list], list]   # exponents -> [(p, is_prime, seconds, res64)]


def modules():
    """(lucaslehmer, lucaslehmer_cuda): the two extension modules built beside this file."""
    from . import add_dropin_path
    add_dropin_path()
    import lucaslehmer
    import lucaslehmer_cuda
    return lucaslehmer, lucaslehmer_cuda


def chunked(exps: list, size: int) -> list:
    """The ascending exponent list cut into calls of at most `size` exponents (0: one call)."""
    if not exps:
        return []
    step = size if size > 0 else len(exps)
    return [exps[i:i + step] for i in range(0, len(exps), step)]


def device_engine(engine: str) -> Engine:
    """A callable list -> rows that runs one device call per list (input order kept)."""
    ll, lc = modules()
    transform = "fp64" if engine == "fft" else "ntt"

    def run(ps: list) -> list:
        rows = []
        for p, prime, secs, res in lc.lucas_lehmer_batch(ps, 0, transform):
            if secs < 0:                       # handed back: redo alone, exactly
                t0 = time.perf_counter()
                prime, res = ll.lucas_lehmer_fft_res(p)
                secs = time.perf_counter() - t0
            rows.append((int(p), bool(prime), float(secs), int(res)))
        return rows
    return run


def table_row(p: int, prime: bool, secs: float, at: float, res: int, digits: Callable[[int], int]) -> str:
    mark = "yes" if p in KNOWN else ("NEW?!" if prime else "")
    return (f"{p:>12} | {('YES' if prime else 'no'):^13} | {secs:>10.4f} | {at:>12.3f} | {res:016x} | "
            f"{digits(p) if prime else '':>8} | {mark}")


HEADER = (f"{'exponent p':>12} | {'2^p-1 prime?':^13} | {'LL time s':>10} | {'from start s':>12} | "
          f"{'res64':>16} | {'digits':>8} | known")


def one_pass(exps: list, args, run: Engine, reference: dict | None, digits, log=print) -> dict:
    """Tests every exponent once; returns the pass record (rows, hits, mismatches, wall)."""
    t0 = time.perf_counter()
    rows, hits, mism = [], [], []
    last = t0
    log(HEADER)
    log("-" * len(HEADER))
    for ps in chunked(sorted(exps), args.cuda_batch):
        out = run(ps)
        at = time.perf_counter() - t0
        lines = []
        for p, prime, secs, res in out:
            rows.append({"p": p, "mersenne": prime, "ll_s": round(secs, 6), "res64": f"{res:016x}"})
            if prime:
                hits.append(p)
            want = None if reference is None else reference.get(p)
            if want is not None and want != f"{res:016x}":
                mism.append(p)
                print(f"RES64 MISMATCH p={p}: got {res:016x}, reference {want}", file=sys.stderr, flush=True)
            if prime or args.all:
                lines.append(table_row(p, prime, secs, at, res, digits))
        if lines:
            log("\n".join(lines))
        now = time.perf_counter()
        if args.progress and now - last >= args.progress:
            print(f"    ... {len(rows)}/{len(exps)} tested, {len(hits)} Mersenne primes, {now - t0:.0f} s, "
                  f"last p={ps[-1]}", file=sys.stderr, flush=True)
            last = now
    return {"rows": rows, "hits": sorted(hits), "mism": mism, "wall": time.perf_counter() - t0}


def parse(argv):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    what = ap.add_mutually_exclusive_group()
    what.add_argument("-r", "--prime_range", type=int, default=0, metavar="HI", help="every prime p in [2, HI]")
    what.add_argument("-p", "--prime_value", type=int, default=0, metavar="P", help="the exponent P alone")
    what.add_argument("-l", "--prime_list", type=int, nargs="+", default=[], metavar="P", help="the exponents given")
    what.add_argument("-n", "--return_num_primes_in_range", type=int, nargs=2, metavar=("LO", "HI"),
                      help="print how many primes lie in [LO, HI] and stop")
    ap.add_argument("-t", "--threads", type=int, default=1,
                    help="host threads (1..256); the device serialises its calls, so one is enough")
    ap.add_argument("--engine", choices=("fft", "gmp", "cuda"), default="fft",
                    help="fft: FP64 IBDWT (default); gmp, cuda: the exact Goldilocks NTT")
    ap.add_argument("--cuda-batch", type=int, default=None, metavar="N",
                    help="exponents per device call (default: the whole list; cuda: 2048)")
    ap.add_argument("--all", action="store_true", help="a table row for every exponent, not only the primes")
    ap.add_argument("--json", metavar="FILE", help="write the rows and timings of the last pass to FILE")
    ap.add_argument("--verify", metavar="FILE", help="check every Res64 against the rows of an earlier --json FILE")
    ap.add_argument("--loop", type=int, default=1, metavar="N",
                    help="repeat the whole job N times (0: until interrupted); later passes are checked "
                         "against the first (or against --verify)")
    ap.add_argument("--progress", type=float, default=10.0, metavar="SEC",
                    help="a progress line on stderr at most every SEC seconds (0: none)")
    a = ap.parse_args(argv)
    if a.cuda_batch is None:
        a.cuda_batch = 2048 if a.engine == "cuda" else 0
    return a


def main(argv: Iterable[str] | None = None, engine: Engine | None = None) -> int:
    """Runs the driver.  `engine` (a callable list -> rows) replaces the device, for host tests."""
    a = parse(list(argv) if argv is not None else sys.argv[1:])
    if a.threads < 1 or a.threads > MAX_WORKERS:
        print(f"workers must be 1..{MAX_WORKERS}", file=sys.stderr)
        return 1
    ll, _ = modules()
    primes, digits = ll.primes, ll.mersenne_digits
    if a.return_num_primes_in_range:
        lo, hi = a.return_num_primes_in_range
        print(f"primes in [{lo}, {hi}]: {len(primes(lo, hi))}")
        return 0
    exps = [a.prime_value] if a.prime_value else list(a.prime_list) or list(primes(2, a.prime_range or 10001))
    run = engine or device_engine(a.engine)

    reference = None
    if a.verify:
        with open(a.verify) as f:
            reference = {int(r["p"]): r.get("res64") for r in json.load(f)["rows"]}
    gil = getattr(sys, "_is_gil_enabled", lambda: True)()
    print(f"python {sys.version.split()[0]}  gil={'on' if gil else 'off'}  engine={a.engine}  "
          f"exponents={len(exps)}  max p={max(exps)}  per call={a.cuda_batch or len(exps)}", flush=True)
    known_in_range = [p for p in KNOWN if min(exps) <= p <= max(exps)]
    passes = 0
    while True:
        passes += 1
        if a.loop != 1:
            src = "checked against " + (a.verify if a.verify else "pass 1") if reference is not None else "makes the reference"
            print(f"--- pass {passes}: {src}", file=sys.stderr, flush=True)
        rec = one_pass(exps, a, run, reference, digits)
        print("-" * len(HEADER))
        print(f"tested {len(exps)} exponents in {rec['wall']:.3f} s wall, "
              f"{sum(r['ll_s'] for r in rec['rows']):.3f} s of LL work, "
              f"{len(rec['hits'])} Mersenne primes: {rec['hits']}", flush=True)
        missing = [p for p in known_in_range if p not in rec["hits"]]
        if missing:
            print(f"WARNING: known Mersenne exponents in range NOT found: {missing}", file=sys.stderr)
        if reference is not None:
            seen = sum(1 for r in rec["rows"] if r["p"] in reference)
            print(f"res64 verify: {seen} checked, {len(rec['mism'])} mismatches"
                  + (f" -> {rec['mism'][:20]}" if rec["mism"] else " -> all match"), file=sys.stderr, flush=True)
        if a.json:
            with open(a.json, "w") as f:
                json.dump({"python": sys.version.split()[0], "gil": gil, "gmp": ll.gmp_version, "workers": a.threads,
                           "executor": "device", "engine": a.engine, "fft_backend": ll.fft_backend, "split": 0,
                           "order": "sorted", "wall_s": round(rec["wall"], 3), "hits": rec["hits"],
                           "rows": rec["rows"]}, f, indent=1)
        if rec["mism"]:
            if a.loop != 1:
                print(f"--- pass {passes}: Res64 mismatch, stopping", file=sys.stderr)
            return 3
        if a.loop != 1 and reference is None:
            reference = {r["p"]: r["res64"] for r in rec["rows"]}
        if a.loop == 1 or (a.loop > 1 and passes >= a.loop):
            return 2 if missing else 0


if __name__ == "__main__":
    sys.exit(main())
