#!/usr/bin/env python3
"""Headline benchmark: Lucas-Lehmer iterations/s of the batched GPU engine.

Workload (one "step"): the full Lucas-Lehmer test of every prime 2 <= p <= 100001
(9592 exponents, sum(p-2) = 454,377,353 squarings) in ONE persistent launch of the
batched FP64 IBDWT engine, every Res64 checked against the committed GMP fixture.

Prints one JSON line (value = iterations/s of the whole step, inputs on the device: the
only input is the exponent list; residues are created and kept on the GPU), with:
  roofline      FP64 VALU bound of the batch kernel: algorithmic flops per launch
                (sum over exponents of (p-2) * (5 N log2 N + 6.5 N)) / kernel time
                (HIP events on the library's stream), against MI355X's 78.6 TFLOP/s FP64;
                traffic from the committed rocprofv3 PMC summary when present.
  cpu_baseline  the CPU oracle's GMP engine (a restatement of lucaslehmer.cpp:29-56, one
                exponent per thread as PerfNumMultiCLL.py:189-193) on a bounded sample.

The library call is synchronous (it waits on its own HIP stream), so the timed region is
bracketed by completed device work on both sides.  One GPU only: --gpus > 1 exits.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def load_fixture(name: str):
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return None
    out = {}
    with open(path) as f:
        next(f)
        for line in f:
            p, v, r = line.strip().split(",")
            out[int(p)] = (int(v), int(r, 16))
    return out


def fp64_flops_per_iter(n: int) -> float:
    return 5.0 * n * math.log2(n) + 6.5 * n


def cpu_baseline(ps: list[int], every: int, threads: int) -> dict:
    from oracle import oracle as O
    sample = ps[every // 2::every]
    iters = sum(p - 2 for p in sample if p > 2)
    t0 = time.perf_counter()
    O.ll_many(sample, engine=0, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "LL iters/s", "cores": threads, "kind": "port",
            "sample": f"GMP engine (restates lucaslehmer.cpp:29-56), full LL of every {every}th prime "
                      f"<= 100001 ({len(sample)} exponents, {iters} squarings, {dt:.1f} s wall), "
                      f"one exponent per thread, GMP {O.gmp_version()}"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--transform", choices=["fp64", "ntt"], default="fp64")
    ap.add_argument("--limit", type=int, default=100001, help="largest exponent of the workload")
    ap.add_argument("--cpu-every", type=int, default=200, help="CPU baseline sample: every k-th prime")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {a.gpus}); the exponents of a batch are "
              "independent, so N GPUs would each take a longest-first share of the list.", file=sys.stderr)
        return 2

    from mersenneprime_python_cpp_gmp_benchmark_amd import llhip
    from mersenneprime_python_cpp_gmp_benchmark_amd import add_dropin_path
    add_dropin_path()
    import lucaslehmer as ll

    tr = llhip.FP64 if a.transform == "fp64" else llhip.NTT
    dev = llhip.device_info()
    ps = ll.primes(2, a.limit)
    total_iters = sum(p - 2 for p in ps if p > 2)
    log(f"device {dev['name']} ({dev['arch']}, {dev['cus']} CUs); {len(ps)} exponents, {total_iters} squarings")

    for w in range(a.warmup):
        t0 = time.perf_counter()
        llhip.batch(ps, 0, tr)
        log(f"warmup {w}: {time.perf_counter() - t0:.2f} s")

    kernel_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for s in range(a.steps):
        out = llhip.batch(ps, 0, tr)
        ms, nl = llhip.last_launch_stats()
        kernel_ms += ms
        launches += nl
    t1 = time.perf_counter()
    wall = t1 - t0
    log(f"timed {a.steps} step(s): {wall:.3f} s, kernel {kernel_ms:.1f} ms over {launches} launch(es)")

    fx = load_fixture("c3_gmp_p100001.csv")
    diffs = None
    if fx is not None:
        diffs = 0
        for p, v, r in zip(ps, out["verdict"], out["res64"]):
            if p in fx and (fx[p][1] != int(r) or fx[p][0] != int(v != 0)):
                diffs += 1
    bad_status = int(sum(1 for s in out["status"] if s not in (llhip.ST_OK, llhip.ST_TRIVIAL)))

    n_used = out["n_used"]
    flops = 0.0
    for p, n in zip(ps, n_used):
        if p > 2 and n > 1:
            flops += (p - 2) * fp64_flops_per_iter(int(n))
    per_launch_s = kernel_ms / 1e3 / max(a.steps, 1)
    achieved = flops / per_launch_s / 1e12 if per_launch_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_batch_fp64_p100001.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": "ll_batch_kernel<FP64>" if tr == llhip.FP64 else "ll_batch_kernel<NTT>",
                "kernel_ms_per_launch": round(per_launch_s * 1e3, 3),
                "flops_per_launch": flops if tr == llhip.FP64 else None}

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(ps, a.cpu_every, a.cpu_threads)
        log(f"cpu baseline: {cpu['value']:.0f} iters/s on {cpu['cores']} threads")

    value = a.steps * total_iters / wall
    line = {
        "metric": "LL iters/s: batched p<=100001 and p~2^24 learned-N vs plain GPU, 0 Res64 diffs",
        "value": round(value, 1), "unit": "LL iters/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if tr == llhip.FP64 else "u64 (GF(2^64-2^32+1))",
        "data": "synthetic (deterministic: the exponent list; s_0 = 4)",
        "config": {"workload": f"batched full Lucas-Lehmer of all {len(ps)} primes p <= {a.limit} in one "
                               f"persistent launch ({a.transform} IBDWT)",
                   "exponents": len(ps), "squarings": total_iters, "transform": a.transform,
                   "bits_per_word_cap": llhip.get_bpw()},
        "res64_diffs_vs_gmp": diffs, "not_ok_status": bad_status,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
