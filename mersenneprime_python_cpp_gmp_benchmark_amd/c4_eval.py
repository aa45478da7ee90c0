"""Learned vs conservative FFT length on the held-out exponents (GPU).

For each held-out exponent of tests/golden/c4_exponents.csv: run K squarings (default 20000)
at the reference's conservative length and at the learned length, both with the per-iteration
round-off check and rollback; compare the Res64 of the two runs with each other and with the
committed GMP fixture where there is one; report the total device time of each.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.c4_eval [--iters K] [--limit L] [--out F.json]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import llhip, nselect
from .roundoff_data import ROOT, exponents


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=20000)
    ap.add_argument("--limit", type=int, default=0)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    ps = exponents("heldout")
    if a.limit:
        ps = ps[:: max(1, len(ps) // a.limit)][: a.limit]
    gmp = {}
    path = os.path.join(ROOT, "tests", "golden", "c4_gmp_heldout.csv")
    if os.path.exists(path):
        for line in list(open(path))[1:]:
            p, k, r = line.strip().split(",")
            gmp[(int(p), int(k))] = int(r, 16)
    model = nselect.load_model()
    rows = []
    tc = tl = 0.0
    for i, p in enumerate(ps):
        nc = nselect.conservative_length(p)
        nl = nselect.choose_length(p, model)
        c = llhip.lucas_lehmer_n(p, a.iters, llhip.FP64, nc)
        l = c if nl == nc else llhip.lucas_lehmer_n(p, a.iters, llhip.FP64, nl)
        want = gmp.get((p, a.iters))
        row = {"p": p, "n_conservative": nc, "n_learned": nl, "n_learned_used": l["n_used"],
               "s_conservative": c["seconds"], "s_learned": l["seconds"], "maxerr_learned": l["maxerr"],
               "res64": f"{l['res64']:016x}", "res64_equal": c["res64"] == l["res64"],
               "gmp": None if want is None else want == l["res64"]}
        rows.append(row)
        tc += c["seconds"]
        tl += l["seconds"]
        print(f"{i + 1}/{len(ps)} p={p} N {nc} -> {nl} ({l['n_used']}) t {c['seconds']:.3f} -> {l['seconds']:.3f} s "
              f"eq={row['res64_equal']} gmp={row['gmp']} err={l['maxerr']:.3g}", flush=True)
    summary = {"exponents": len(rows), "iters": a.iters, "seconds_conservative": tc, "seconds_learned": tl,
               "speedup": tc / tl if tl else None, "shorter": sum(r["n_learned"] < r["n_conservative"] for r in rows),
               "res64_mismatch_learned_vs_conservative": sum(not r["res64_equal"] for r in rows),
               "gmp_checked": sum(r["gmp"] is not None for r in rows),
               "gmp_mismatch": sum(r["gmp"] is False for r in rows),
               "rollbacks": sum(r["n_learned_used"] != r["n_learned"] for r in rows)}
    print(json.dumps(summary), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump({"summary": summary, "rows": rows}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
