"""Learned vs conservative FFT length on the held-out exponents (GPU).

The held-out exponents of tests/golden/c4_exponents.csv (primes drawn with seed 0 from
[2^20, 2^25), every 4th held out of the round-off model's training) each run K squarings
(default 20000) with the per-iteration round-off check and rollback:

  plain:        one exponent per call at the reference's conservative length (the plain
                GPU form);
  conservative: all exponents in one call at the conservative lengths (the large engine keeps
                several exponents in flight);
  learned:      all exponents in one call at the lengths nselect.choose_length picks with the
                costs measured with several exponents in flight.

Res64 of every run is compared with the others and with the committed GMP fixture where there
is one; the device time of each is reported.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.c4_eval [--iters K] [--limit L] [--out F.json]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import llhip, nselect
from .roundoff_data import ROOT, exponents


def gmp_fixture() -> dict:
    out = {}
    path = os.path.join(ROOT, "tests", "golden", "c4_gmp_heldout.csv")
    if os.path.exists(path):
        for line in list(open(path))[1:]:
            p, k, r = line.strip().split(",")
            out[(int(p), int(k))] = int(r, 16)
    return out


def compare(iters: int = 20000, limit: int = 0, plain: bool = True, log=print) -> dict:
    """Run the three forms over the held-out exponents; returns {"summary", "rows"}."""
    ps = exponents("heldout")
    if limit:
        ps = ps[:: max(1, len(ps) // limit)][:limit]
    gmp = gmp_fixture()
    model = nselect.load_model()
    nc = [nselect.conservative_length(p) for p in ps]
    nl = [nselect.choose_length(p, model, loaded=True) for p in ps]
    t_plain = None
    single = None
    if plain:
        t0 = time.perf_counter()
        single = [llhip.lucas_lehmer_n(p, iters, llhip.FP64, n) for p, n in zip(ps, nc)]
        t_plain = time.perf_counter() - t0
        log(f"plain (one exponent per call, conservative length): {t_plain:.2f} s")
    t0 = time.perf_counter()
    cons = llhip.batch_n(ps, iters, llhip.FP64, nc)
    t_cons = time.perf_counter() - t0
    log(f"one call, conservative lengths: {t_cons:.2f} s")
    t0 = time.perf_counter()
    lrn = llhip.batch_n(ps, iters, llhip.FP64, nl)
    t_lrn = time.perf_counter() - t0
    log(f"one call, learned lengths: {t_lrn:.2f} s")
    rows = []
    for i, p in enumerate(ps):
        want = gmp.get((p, iters))
        r = {"p": p, "n_conservative": nc[i], "n_learned": nl[i], "n_learned_used": int(lrn["n_used"][i]),
             "maxerr_learned": float(lrn["maxerr"][i]), "res64": f"{int(lrn['res64'][i]):016x}",
             "res64_equal": int(lrn["res64"][i]) == int(cons["res64"][i])
             and (single is None or single[i]["res64"] == int(lrn["res64"][i])),
             "gmp": None if want is None else want == int(lrn["res64"][i])}
        rows.append(r)
    ok = (np.asarray(lrn["status"]) == llhip.ST_OK).all() and (np.asarray(cons["status"]) == llhip.ST_OK).all()
    summary = {"exponents": len(rows), "iters": iters, "all_ok": bool(ok),
               "seconds_plain": None if t_plain is None else round(t_plain, 3),
               "seconds_one_call_conservative": round(t_cons, 3), "seconds_one_call_learned": round(t_lrn, 3),
               "speedup_learned_vs_plain": None if t_plain is None else round(t_plain / t_lrn, 3),
               "speedup_learned_vs_conservative": round(t_cons / t_lrn, 3),
               "shorter": sum(r["n_learned"] < r["n_conservative"] for r in rows),
               "res64_mismatch": sum(not r["res64_equal"] for r in rows),
               "gmp_checked": sum(r["gmp"] is not None for r in rows),
               "gmp_mismatch": sum(r["gmp"] is False for r in rows),
               "rollbacks": sum(r["n_learned_used"] != r["n_learned"] for r in rows)}
    return {"summary": summary, "rows": rows}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=20000)
    ap.add_argument("--limit", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true", help="skip the one-exponent-per-call run")
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    res = compare(a.iters, a.limit, not a.no_plain, log=lambda m: print(m, flush=True))
    print(json.dumps(res["summary"]), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
