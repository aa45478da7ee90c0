#!/usr/bin/env python3
"""Writes the committed golden fixtures from the CPU oracle's GMP engine.

The GMP engine restates lucaslehmer.cpp:29-56 (the reference's exact arbiter); Res64 is
the low 64 bits of the canonical residue.  Each fixture is plain CSV:

  c1_gmp_p10001.csv      every prime 2 <= p <= 10001, full test: p,prime,res64
  c3_gmp_p100001.csv     every prime 2 <= p <= 100001, full test (hours of CPU)
  big_gmp_interim.csv    interim residues: p,iters,res64 (p = 1,048,573 @ 10,000;
                         p = 16,777,213 @ 5,000; p = 3,021,377 @ 20,000 = README.md:200-206)
  c4_exponents.csv       256 primes from [2^20, 2^25), numpy default_rng(0), with split
  c4_gmp_heldout.csv     the 64 held-out exponents: p,iters,res64 at 20,000 iterations
  rollback_gmp.csv       p,iters,res64 of the large engine's rollback cases at 6,000 squarings
  above_batch_gmp.csv    the four primes just above 21 x 8192 (beyond the batched FP64
                         lengths, so a batch call runs them on the large engine), full test:
                         p,prime,res64
  concurrent_gmp.csv     the ten exponents of the concurrent large-job test, p,iters,res64 at 400
                         squarings
  sched_gmp.csv          one exponent per large-engine length N (the largest prime <= 16.5 N),
                         p,iters,res64 at 201 squarings (two graph blocks: the block boundary's
                         carry buffer is crossed once)

Usage:  python tests/golden/make_fixtures.py c1|c3|big|c4list|c4gmp|above|rollback|sched|concurrent [--threads T]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def run_range(hi: int, path: str, threads: int) -> None:
    ps = O.primes(2, hi)
    # longest first so the threads finish together; written in ascending order
    order = sorted(ps, reverse=True)
    done = {}
    if os.path.exists(path + ".partial"):
        for line in open(path + ".partial"):
            p, v, r = line.strip().split(",")
            done[int(p)] = (int(v), r)
    todo = [p for p in order if p not in done]
    t0 = time.time()
    chunk = max(threads * 4, 16)
    with open(path + ".partial", "a") as f:
        for i in range(0, len(todo), chunk):
            part = todo[i:i + chunk]
            out = O.ll_many(part, engine=0, threads=threads)
            for p, (v, r) in zip(part, out):
                f.write(f"{p},{int(v)},{r:016x}\n")
                done[p] = (int(v), f"{r:016x}")
            f.flush()
            print(f"{len(done)}/{len(ps)} done, last p={part[-1]}, {time.time() - t0:.0f} s", flush=True)
    with open(path, "w") as f:
        f.write("p,prime,res64\n")
        for p in ps:
            v, r = done[p]
            f.write(f"{p},{v},{r}\n")
    os.remove(path + ".partial")


def run_big(path: str) -> None:
    cases = [(3021377, 20000), (1048573, 10000), (16777213, 5000)]
    rows = []
    for p, k in cases:
        t0 = time.time()
        v, r, _ = O.ll_gmp(p, k)
        rows.append((p, k, r))
        print(f"p={p} k={k} res64={r:016x} {time.time() - t0:.1f} s", flush=True)
    assert rows[0][2] == 0xCE811E3772129824, "README.md:200-206 interim residue mismatch"
    with open(path, "w") as f:
        f.write("p,iters,res64\n")
        for p, k, r in rows:
            f.write(f"{p},{k},{r:016x}\n")


def c4_list(path: str) -> None:
    import numpy as np
    ps = O.primes(1 << 20, (1 << 25) - 1)
    rng = np.random.default_rng(0)
    pick = sorted(int(x) for x in rng.choice(np.array(ps, dtype=np.int64), 256, replace=False))
    # fixed split: every 4th (by ascending p) is held out -> 64 held-out, 192 train
    with open(path, "w") as f:
        f.write("p,split\n")
        for i, p in enumerate(pick):
            f.write(f"{p},{'heldout' if i % 4 == 3 else 'train'}\n")
    print(f"{len(ps)} primes in [2^20, 2^25); wrote {len(pick)}")


def _read_rows(path: str) -> dict:
    """{p: res64 hex} of a p,iters,res64 file (header optional); {} if absent."""
    out = {}
    if os.path.exists(path):
        for line in open(path):
            a = line.strip().split(",")
            if len(a) == 3 and a[0].isdigit():
                out[int(a[0])] = a[2]
    return out


def c4_gmp(list_path: str, path: str, threads: int, count: int, order: str, partial: str) -> None:
    """GMP residues of the held-out exponents after 20,000 squarings, one exponent per thread.

    Resumable: every finished exponent is appended to `partial` at once, and exponents already
    in `path` or in any c4_gmp_heldout*.partial file beside it are skipped.  `order` asc starts
    with the smallest exponents, desc with the largest (two machines can split the list).  When
    every held-out exponent has a residue, `path` is rewritten sorted by p."""
    import glob
    from concurrent.futures import ThreadPoolExecutor, as_completed
    rows = [l.strip().split(",") for l in open(list_path)][1:]
    held = [int(p) for p, s in rows if s == "heldout"]
    done = _read_rows(path)
    for f in glob.glob(os.path.join(os.path.dirname(path), "c4_gmp_heldout*.partial")) + [partial]:
        done.update(_read_rows(f))
    todo = sorted((p for p in held if p not in done), reverse=order == "desc")[:count]
    t0 = time.time()
    print(f"{len(done)} held-out residues known, {len(todo)} to run on {threads} threads", flush=True)
    with open(partial, "a") as f, ThreadPoolExecutor(max_workers=threads) as ex:
        futs = {ex.submit(O.ll_gmp, p, 20000): p for p in todo}   # ctypes drops the GIL
        for fu in as_completed(futs):
            p = futs[fu]
            _, r, _ = fu.result()
            f.write(f"{p},20000,{r:016x}\n")
            f.flush()
            done[p] = f"{r:016x}"
            print(f"p={p} res64={r:016x} ({time.time() - t0:.0f} s)", flush=True)
    if all(p in done for p in held):
        with open(path, "w") as f:
            f.write("p,iters,res64\n")
            for p in held:
                f.write(f"{p},20000,{done[p]}\n")
        print(f"all {len(held)} held-out residues written to {path}", flush=True)


def rollback(path: str, threads: int) -> None:
    """GMP residues of the large engine's rollback cases after 6,000 squarings: exponents whose
    FP64 round-off first exceeds 0.4 mid-run at a forced length (p = 1,333,663 at N = 65,536,
    p = 1,658,893 at N = 81,920)."""
    cases = [1333663, 1658893]
    out = O.ll_many(cases, iters=6000, engine=0, threads=threads)
    with open(path, "w") as f:
        f.write("p,iters,res64\n")
        for p, (_, r) in zip(cases, out):
            f.write(f"{p},6000,{r:016x}\n")
    print(f"{len(cases)} rollback residues written")


def above_batch(path: str, threads: int) -> None:
    ps = O.primes(21 * 8192 + 1, 21 * 8192 + 200)[:4]
    t0 = time.time()
    out = O.ll_many(ps, engine=0, threads=threads)
    with open(path, "w") as f:
        f.write("p,prime,res64\n")
        for p, (v, r) in zip(ps, out):
            f.write(f"{p},{int(v)},{r:016x}\n")
    print(f"{len(ps)} exponents in {time.time() - t0:.0f} s")


# Large-engine lengths N (digit counts) of both transforms: FP64 N = 2M for M in
# {1, 3, 5, 7, 15} * 2^k, 2^15 <= M <= 2^21 (M >= 5 * 2^13 for the odd factors); NTT N = M = 2^15 .. 2^21.
SCHED_ITERS = 201


def sched_lengths() -> list:
    fp64 = [m << k for m in (1, 3, 5, 7, 15) for k in range(13, 22) if (1 << 15) <= (m << k) <= (1 << 21)
            and (m == 1 or (m << k) >= (5 << 13))]
    return sorted(set(2 * m for m in fp64) | set(1 << k for k in range(15, 22)))


def sched_exponent(n: int) -> int:
    hi = int(16.5 * n)
    return O.primes(hi - 20000, hi)[-1]


def sched(path: str, threads: int) -> None:
    ps = sorted(set(sched_exponent(n) for n in sched_lengths()), reverse=True)
    t0 = time.time()
    out = O.ll_many(ps, iters=SCHED_ITERS, engine=0, threads=threads)
    with open(path, "w") as f:
        f.write("p,iters,res64\n")
        for p, (_, r) in sorted(zip(ps, out)):
            f.write(f"{p},{SCHED_ITERS},{r:016x}\n")
    print(f"{len(ps)} schedule-test residues in {time.time() - t0:.0f} s")


CONCURRENT = [1048573, 1299709, 2000003, 2750159, 4194301, 5000011, 6291469, 8388593, 1572869, 3145739]


def concurrent(path: str, threads: int) -> None:
    out = O.ll_many(CONCURRENT, iters=400, engine=0, threads=threads)
    with open(path, "w") as f:
        f.write("p,iters,res64\n")
        for p, (_, r) in zip(CONCURRENT, out):
            f.write(f"{p},400,{r:016x}\n")
    print(f"{len(CONCURRENT)} concurrent-test residues written")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["c1", "c3", "big", "c4list", "c4gmp", "above", "rollback", "sched",
                                    "concurrent"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--order", choices=["asc", "desc"], default="asc")
    ap.add_argument("--partial", default=os.path.join(HERE, "c4_gmp_heldout.partial"))
    a = ap.parse_args()
    if a.what == "c1":
        run_range(10001, os.path.join(HERE, "c1_gmp_p10001.csv"), a.threads)
    elif a.what == "c3":
        run_range(100001, os.path.join(HERE, "c3_gmp_p100001.csv"), a.threads)
    elif a.what == "big":
        run_big(os.path.join(HERE, "big_gmp_interim.csv"))
    elif a.what == "above":
        above_batch(os.path.join(HERE, "above_batch_gmp.csv"), a.threads)
    elif a.what == "rollback":
        rollback(os.path.join(HERE, "rollback_gmp.csv"), a.threads)
    elif a.what == "concurrent":
        concurrent(os.path.join(HERE, "concurrent_gmp.csv"), a.threads)
    elif a.what == "sched":
        sched(os.path.join(HERE, "sched_gmp.csv"), a.threads)
    elif a.what == "c4list":
        c4_list(os.path.join(HERE, "c4_exponents.csv"))
    else:
        c4_gmp(os.path.join(HERE, "c4_exponents.csv"), os.path.join(HERE, "c4_gmp_heldout.csv"),
               a.threads, a.count, a.order, a.partial)


if __name__ == "__main__":
    main()
