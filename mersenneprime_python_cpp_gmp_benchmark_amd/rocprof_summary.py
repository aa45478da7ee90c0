"""Summarise a rocprofv3 results database (``*_results.db``) as a short markdown table.

    python -m mersenneprime_python_cpp_gmp_benchmark_amd.rocprof_summary DB "COMMAND" [OUT.md]

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


def main() -> int:
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
