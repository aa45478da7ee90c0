"""Range driver: the counterpart of the reference's PerfNumMultiCLL.py (same options, output table,
JSON rows, --verify / --loop torture mode and chunked batch dispatch), every Lucas-Lehmer test on
the GPU.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -r 100001            # every prime p <= 100001
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -p 11213              # one exponent
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -l 521 607 1279       # a list
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -n 2 10001            # count primes
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -r 10001 --json a.json
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.driver -r 10001 --verify a.json --loop 3

Engines (--engine):
  hip (default)  ascending chunks of --batch exponents (0 = all at once) through
                 lucaslehmer_hip.lucas_lehmer_batch: persistent batched launches up to the batch
                 limit, concurrent large-engine jobs above it; FP64 IBDWT with per-squaring round-off
                 check and rollback (--transform ntt for the exact Goldilocks NTT);
  cuda           the reference's name for its GPU engine: the same batches on the exact NTT;
  fft            one lucaslehmer.lucas_lehmer_fft_res call per exponent from a pool of workers
                 (PerfNumMultiCLL.py:33-44): the FP64 engine, redone exactly on rounding trouble;
  gmp            one lucaslehmer.lucas_lehmer_res call per exponent: the exact engine.
Every engine runs on the device.  Where the reference falls back to its CPU FFT engine for an
exponent the GPU kernel cannot take (PerfNumMultiCLL.py:148-151), this driver redoes it with the
device FP64 engine; nothing runs on the CPU.  --split is accepted and has no effect (it sets the
thread split of the reference's CPU GMP engine).

Exit codes as the reference: 0 ok, 1 bad arguments, 2 a known Mersenne exponent in range was not
found, 3 a Res64 mismatch under --verify.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing
import os
import sys
import tempfile
import time
from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor, as_completed

MAX_WORKERS = 256

# Exponents of the known Mersenne primes (public GIMPS record), for the "known" column and the
# end-of-run check of PerfNumMultiCLL.py:226-227.
KNOWN = (2, 3, 5, 7, 13, 17, 19, 31, 61, 89, 107, 127, 521, 607, 1279, 2203, 2281, 3217, 4253, 4423,
         9689, 9941, 11213, 19937, 21701, 23209, 44497, 86243, 110503, 132049, 216091, 756839, 859433,
         1257787, 1398269, 2976221, 3021377, 6972593, 13466917, 20996011, 24036583, 25964951, 30402457,
         32582657, 37156667, 42643801, 43112609, 57885161, 74207281, 77232917, 82589933, 136279841)


def _ll():
    from . import add_dropin_path       # the package directory holds the two extension modules
    add_dropin_path()
    import lucaslehmer
    return lucaslehmer


def _llhip():
    from . import add_dropin_path
    add_dropin_path()
    import lucaslehmer_hip
    return lucaslehmer_hip


def test_one(ll, engine: str, p: int):
    """(p, is_prime, seconds, res64) of one exponent (PerfNumMultiCLL.py:33-44)."""
    t0 = time.perf_counter()
    if engine == "fft":
        r, res = ll.lucas_lehmer_fft_res(p)
    else:
        r, res = ll.lucas_lehmer_res(p)
    return p, bool(r), time.perf_counter() - t0, int(res)


def _worker_test(engine: str, p: int):          # process-pool entry point
    return test_one(_ll(), engine, p)


def batch_chunks(exps_sorted: list[int], size: int) -> list[list[int]]:
    """Ascending chunks of at most `size` exponents (0 = one chunk): one launch each, so sorted
    output streams chunk by chunk (PerfNumMultiCLL.py:141-146; the batched engine takes uneven
    lengths in one launch, so chunks need not share a length)."""
    if size <= 0:
        return [list(exps_sorted)] if exps_sorted else []
    return [list(exps_sorted[i:i + size]) for i in range(0, len(exps_sorted), size)]


def load_reference(path: str) -> dict[int, str]:
    with open(path) as f:
        return {int(row["p"]): row.get("res64") for row in json.load(f)["rows"]}


def _strip_loop(argv: list[str]) -> list[str]:
    base = list(argv)
    for i, x in enumerate(base):
        if x == "--loop":
            del base[i:i + 2]
            break
        if x.startswith("--loop="):
            del base[i]
            break
    return base


def _strip_opt(argv: list[str], name: str) -> list[str]:
    out, skip = [], False
    for x in argv:
        if skip:
            skip = False
            continue
        if x == name:
            skip = True
            continue
        if x.startswith(name + "="):
            continue
        out.append(x)
    return out


def parse(argv):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("-t", "--threads", type=int, default=min(10, os.cpu_count() or 1),
                    help="number of workers for the per-exponent engines (default 10 or cpu count)")
    ap.add_argument("--workers", choices=("threads", "processes"), default="threads", help="executor kind")
    g = ap.add_mutually_exclusive_group()
    g.add_argument("-r", "--prime_range", type=int, default=0, help="test every prime exponent p from 2 to this number")
    g.add_argument("-p", "--prime_value", type=int, default=0, help="test one exponent")
    g.add_argument("-l", "--prime_list", type=int, nargs="+", default=[], help="test these exponents")
    g.add_argument("-n", "--return_num_primes_in_range", type=int, nargs=2, metavar=("LO", "HI"),
                   help="just count the primes in [LO, HI]")
    ap.add_argument("--all", action="store_true", help="print every tested exponent, not only the Mersenne hits")
    ap.add_argument("--json", metavar="FILE", help="write results + timings to FILE")
    ap.add_argument("--progress", type=float, default=10.0, metavar="SEC", help="progress line every SEC s (0 = off)")
    ap.add_argument("--order", choices=("sorted", "done"), default="sorted",
                    help="sorted: ascending p as soon as every smaller exponent finished; done: completion order")
    ap.add_argument("--engine", choices=("hip", "cuda", "fft", "gmp"), default="hip", help="see the module docstring")
    ap.add_argument("--transform", choices=("fp64", "ntt"), default="fp64", help="hip engine transform (cuda: ntt)")
    ap.add_argument("--verify", metavar="RUN.json", help="compare every res64 against a previous --json run (exit 3 on mismatch)")
    ap.add_argument("--loop", type=int, default=1, metavar="N",
                    help="run the whole job N times (0 = until interrupted); pass 1 is the reference unless --verify")
    ap.add_argument("--batch", "--cuda-batch", dest="batch", type=int, default=0, metavar="N",
                    help="hip/cuda: exponents per launch (default 0: all in one launch)")
    ap.add_argument("--split", type=int, choices=(0, 1, 2), default=0, help="accepted; no effect on the device engines")
    ap.add_argument("--gap", type=float, default=3.0, metavar="SEC",
                    help="sorted mode: flush output blocks SEC s old or 200 lines long (0 = every line)")
    return ap.parse_args(argv)


def main(argv=None, engine_fn=None) -> int:
    """The driver.  engine_fn(engine, exps_sorted, args) -> iterator over chunks of
    [(p, is_prime, seconds, res64)] replaces the device engines (tests of the host logic)."""
    argv = list(argv if argv is not None else sys.argv[1:])
    a = parse(argv)
    if a.loop != 1:                     # torture mode (PerfNumMultiCLL.py:86-101)
        base = _strip_loop(argv)
        ref = a.verify
        tmp = a.json or os.path.join(tempfile.gettempdir(), f"ll_res64_ref_{os.getpid()}.json")
        n = 0
        while a.loop == 0 or n < a.loop:
            n += 1
            args_i = _strip_opt(base, "--verify")
            if ref:
                args_i += ["--verify", ref]
            if not a.json:
                args_i += ["--json", tmp]
            print(f"=== pass {n}" + (f" (verifying against {ref})" if ref else " (reference pass)"),
                  file=sys.stderr, flush=True)
            rc = main(args_i, engine_fn)
            if rc == 3:
                print(f"=== pass {n}: RES64 MISMATCH -- stopping", file=sys.stderr)
                return 3
            if not ref:
                ref = a.json or tmp
        return 0

    ll = _ll()
    if a.return_num_primes_in_range:
        lo, hi = a.return_num_primes_in_range
        print(f"primes in [{lo}, {hi}]: {len(ll.primes(lo, hi))}")
        return 0
    if a.prime_value:
        exps = [a.prime_value]
    elif a.prime_list:
        exps = list(a.prime_list)
    else:
        exps = list(ll.primes(2, a.prime_range or 10001))
    if a.threads < 1 or a.threads > MAX_WORKERS:
        print(f"workers must be 1..{MAX_WORKERS}", file=sys.stderr)
        return 1
    if a.split:
        print("note: --split sets the reference's CPU GMP thread split; the device engines ignore it", file=sys.stderr)
    engine = a.engine
    transform = "ntt" if engine == "cuda" else a.transform

    known = set(KNOWN)
    gil = getattr(sys, "_is_gil_enabled", lambda: True)()
    print(f"python {sys.version.split()[0]}  gil={'on' if gil else 'off'}  engine={engine}"
          + (f" ({transform})" if engine in ("hip", "cuda") else "")
          + f"  workers={a.threads} ({a.workers})  exponents={len(exps)}  max p={max(exps)}", flush=True)
    hdr = (f"{'exponent p':>12} | {'2^p-1 prime?':^13} | {'LL time s':>10} | {'from start s':>12} | "
           f"{'res64':>16} | {'digits':>8} | known")
    print(hdr)
    print("-" * len(hdr), flush=True)
    t_start = time.perf_counter()
    hits, rows, mism = [], [], []
    state = {"done": 0, "last_prog": t_start}
    ref = load_reference(a.verify) if a.verify else None

    def record(p, r, dt, res):
        state["done"] += 1
        rows.append({"p": p, "mersenne": bool(r), "ll_s": round(dt, 6), "res64": f"{res:016x}"})
        if r:
            hits.append(p)
        if ref is not None and p in ref and ref[p] != f"{res:016x}":
            mism.append(p)
            print(f"RES64 MISMATCH p={p}: got {res:016x}, reference {ref[p]}", file=sys.stderr, flush=True)

    def line(p, r, dt, at, res):
        return (f"{p:>12} | {('YES' if r else 'no'):^13} | {dt:>10.4f} | {at:>12.3f} | {res:016x} | "
                f"{ll.mersenne_digits(p) if r else '':>8} | {'yes' if p in known else ('NEW?!' if r else '')}")

    def progress(extra=""):
        now = time.perf_counter()
        if a.progress and now - state["last_prog"] >= a.progress:
            print(f"    ... {state['done']}/{len(exps)} done, {len(hits)} hits, {now - t_start:.0f} s elapsed{extra}",
                  file=sys.stderr, flush=True)
            state["last_prog"] = now

    ordered = a.order == "sorted"
    exps_sorted = sorted(exps)
    if engine in ("hip", "cuda") or engine_fn is not None:
        if engine_fn is None:
            lh = _llhip()
            name, cus, _mem = lh.device_info()
            print(f"{engine}: {name}, {cus} CUs, batched up to p = {lh.max_exponent()} (NTT), larger p on the "
                  f"large engine", file=sys.stderr, flush=True)

            def run_chunks():
                for ps in batch_chunks(exps_sorted, a.batch):
                    got = {int(p): (bool(r), float(dt), int(rs)) for p, r, dt, rs in lh.lucas_lehmer_batch(ps, 0, transform)}
                    out = []
                    for p in ps:
                        r, dt, rs = got.get(p, (False, -1.0, 0))
                        if dt < 0:              # not taken by the batch: the device FP64 engine, exact redo
                            _, r, dt, rs = test_one(ll, "fft", p)
                        out.append((p, r, dt, rs))
                    yield out
            chunks = run_chunks()
        else:
            chunks = engine_fn(engine, exps_sorted, a)
        for chunk in chunks:
            block = []
            for p, r, dt, rs in chunk:
                at = time.perf_counter() - t_start
                record(p, r, dt, rs)
                if r or a.all:
                    block.append(line(p, r, dt, at, rs))
            if block:
                print("\n".join(block), flush=True)
            progress(f", last p={chunk[-1][0]}" if chunk else "")
    else:
        results = {}
        frontier = 0
        block, block_born = [], None
        submit_order = exps_sorted if ordered else sorted(exps, reverse=True)
        if a.workers == "threads":
            pool = ThreadPoolExecutor(max_workers=a.threads)
        else:
            # fresh interpreters: a forked copy of a process that has used the device is not safe
            pool = ProcessPoolExecutor(max_workers=a.threads, mp_context=multiprocessing.get_context("spawn"))
        with pool as ex:
            if a.workers == "threads":
                futs = [ex.submit(test_one, ll, engine, p) for p in submit_order]
            else:
                futs = [ex.submit(_worker_test, engine, p) for p in submit_order]
            for f in as_completed(futs):
                p, r, dt, rs = f.result()
                now = time.perf_counter()
                at = now - t_start
                record(p, r, dt, rs)
                if not ordered:
                    if r or a.all:
                        print(line(p, r, dt, at, rs), flush=True)
                else:
                    results[p] = (r, dt, at, rs)
                    while frontier < len(exps_sorted) and exps_sorted[frontier] in results:
                        q = exps_sorted[frontier]
                        rq, dq, aq, rsq = results.pop(q)
                        frontier += 1
                        if rq or a.all:
                            if not block:
                                block_born = now
                            block.append(line(q, rq, dq, aq, rsq))
                    if block and (a.gap == 0 or len(block) >= 200):
                        print("\n".join(block), flush=True)
                        block = []
                progress(f", next in order: p={exps_sorted[frontier]}" if ordered and frontier < len(exps_sorted) else "")
                if ordered and block and a.gap and now - block_born >= a.gap:
                    print("\n".join(block), flush=True)
                    block = []
        while ordered and frontier < len(exps_sorted):
            q = exps_sorted[frontier]
            rq, dq, aq, rsq = results.pop(q)
            frontier += 1
            if rq or a.all:
                block.append(line(q, rq, dq, aq, rsq))
        if block:
            print("\n".join(block), flush=True)

    total = time.perf_counter() - t_start
    hits.sort()
    print("-" * len(hdr), flush=True)
    print(f"tested {len(exps)} exponents in {total:.3f} s wall, {sum(r['ll_s'] for r in rows):.3f} s of LL work, "
          f"{len(hits)} Mersenne primes: {hits}")
    missing = [p for p in sorted(known) if min(exps) <= p <= max(exps) and p not in hits]
    if missing:
        print(f"WARNING: known Mersenne exponents in range NOT found: {missing}", file=sys.stderr)
    if ref is not None:
        checked = sum(1 for row in rows if row["p"] in ref)
        print(f"res64 verify against {a.verify}: {checked} checked, {len(mism)} mismatches"
              + (f" -> {mism[:20]}" if mism else " -> all match"), file=sys.stderr, flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"python": sys.version.split()[0], "gil": gil, "gmp": ll.gmp_version, "workers": a.threads,
                       "executor": a.workers, "engine": engine, "transform": transform if engine in ("hip", "cuda") else None,
                       "fft_backend": ll.fft_backend, "split": a.split, "order": a.order, "wall_s": round(total, 3),
                       "hits": hits, "rows": sorted(rows, key=lambda r: r["p"])}, f, indent=1)
    if mism:
        return 3
    return 0 if not missing else 2


if __name__ == "__main__":
    sys.exit(main())
