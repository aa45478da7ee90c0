"""Summarise a rocprofv3 results database (``*_results.db``) as a short markdown table.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.rocprof_summary DB "COMMAND" [OUT.md]
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.rocprof_summary --traffic FETCH_DB WRITE_DB \
        COPY_FETCH_DB COPY_WRITE_DB COPY_BYTES
    python -m mersenneprime_python_cpp_gmp_benchmark_amd.rocprof_summary --counters MATCH DB [DB ...]

Reads the ``kernels`` view (one row per dispatch) and prints, per kernel name: calls,
total / average / min / max duration, grid and workgroup size, VGPR/AGPR/SGPR counts and
LDS bytes -- the figures the roofline of bench.py is checked against.
"""
from __future__ import annotations

import sqlite3
import sys
from collections import defaultdict


def summarise(db: str, command: str) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, duration, grid_x, workgroup_x, vgpr_count, accum_vgpr_count, sgpr_count, "
                     "lds_size from kernels").fetchall()
    agg = defaultdict(list)
    meta = {}
    for name, dur, gx, wx, vg, ag, sg, lds in rows:
        agg[name].append(dur)
        meta[name] = (gx, wx, vg, ag, sg, lds)
    total = sum(sum(v) for v in agg.values()) or 1
    out = [f"command: `{command}`", "",
           "| kernel | calls | total ms | avg ms | min ms | max ms | % | grid | wg | vgpr | agpr | sgpr | lds B |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, durs in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        gx, wx, vg, ag, sg, lds = meta[name]
        short = name if len(name) < 90 else name[:87] + "..."
        out.append(f"| `{short}` | {len(durs)} | {sum(durs) / 1e6:.3f} | {sum(durs) / len(durs) / 1e6:.4f} | "
                   f"{min(durs) / 1e6:.4f} | {max(durs) / 1e6:.4f} | {100 * sum(durs) / total:.2f} | {gx} | {wx} | "
                   f"{vg} | {ag} | {sg} | {lds} |")
    return "\n".join(out) + "\n"


def counter_per_kernel(db: str, counter: str) -> dict:
    """{kernel name: [value per dispatch]} of one PMC counter."""
    c = sqlite3.connect(db)
    out = defaultdict(list)
    for name, cn, v in c.execute("select kernel_name, counter_name, value from counters_collection"):
        if cn == counter:
            out[name].append(float(v))
    return out


def traffic(fetch_db: str, write_db: str, copy_fetch_db: str, copy_write_db: str, copy_bytes: int) -> dict:
    """HBM bytes per dispatch of every kernel from FETCH_SIZE / WRITE_SIZE (KiB per dispatch),
    each scaled by what the counter reads for a device copy of copy_bytes."""
    def copy_kib(db, cn):
        d = counter_per_kernel(db, cn)
        vals = [max(v) for k, v in d.items() if "copy" in k.lower()]
        return max(vals)
    fs = copy_bytes / 1024 / copy_kib(copy_fetch_db, "FETCH_SIZE")
    ws = copy_bytes / 1024 / copy_kib(copy_write_db, "WRITE_SIZE")
    f = counter_per_kernel(fetch_db, "FETCH_SIZE")
    w = counter_per_kernel(write_db, "WRITE_SIZE")
    rows = {}
    for k in f:
        fk = sum(f[k]) / len(f[k])
        wk = sum(w.get(k, [0.0])) / max(1, len(w.get(k, [])))
        rows[k] = {"calls": len(f[k]), "fetch_kib_raw": fk, "write_kib_raw": wk,
                   "bytes": fk * 1024 * fs + wk * 1024 * ws}
    return {"fetch_scale": fs, "write_scale": ws, "kernels": rows}


def counters(dbs: list, match: str = "") -> str:
    """Markdown table: per kernel (name containing `match`), the average per dispatch of every PMC
    counter found in the given databases (one counter set per rocprofv3 --pmc pass)."""
    table = defaultdict(dict)
    calls = {}
    for db in dbs:
        c = sqlite3.connect(db)
        acc = defaultdict(lambda: defaultdict(list))
        for name, cn, v in c.execute("select kernel_name, counter_name, value from counters_collection"):
            if match in name:
                acc[name][cn].append(float(v))
        for name, d in acc.items():
            for cn, vals in d.items():
                table[name][cn] = sum(vals) / len(vals)
                calls[name] = len(vals)
    names = sorted(table, key=lambda n: -calls[n])
    cols = sorted({cn for n in names for cn in table[n]})
    short = [n.replace("void (anonymous namespace)::", "").split("(")[0] for n in names]
    out = ["| counter (average per dispatch) | " + " | ".join(f"`{s}`" for s in short) + " |",
           "|---|" + "---|" * len(names), "| dispatches | " + " | ".join(str(calls[n]) for n in names) + " |"]
    for cn in cols:
        out.append(f"| {cn} | " + " | ".join(f"{table[n].get(cn, float('nan')):.4g}" for n in names) + " |")
    return "\n".join(out) + "\n"


def main() -> int:
    if len(sys.argv) > 1 and sys.argv[1] == "--counters":
        # --counters MATCH DB [DB ...]
        print(counters(sys.argv[3:], sys.argv[2]))
        return 0
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic":
        # --traffic FETCH_DB WRITE_DB COPY_FETCH_DB COPY_WRITE_DB COPY_BYTES
        t = traffic(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], int(sys.argv[6]))
        print(f"fetch scale {t['fetch_scale']:.4f}, write scale {t['write_scale']:.4f}")
        for k, r in sorted(t["kernels"].items(), key=lambda kv: -kv[1]["bytes"]):
            short = k.replace("void (anonymous namespace)::", "").split("(")[0][:80]
            print(f"| `{short}` | {r['calls']} | {r['fetch_kib_raw']:.1f} | {r['write_kib_raw']:.1f} | "
                  f"{r['bytes'] / 1e6:.2f} |")
        return 0
    if len(sys.argv) < 3:
        print(__doc__)
        return 2
    txt = summarise(sys.argv[1], sys.argv[2])
    if len(sys.argv) > 3:
        with open(sys.argv[3], "w") as f:
            f.write(txt)
    print(txt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
