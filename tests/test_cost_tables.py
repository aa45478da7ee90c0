"""The autotuner's tables are measured on the kernels in the tree: each records a hash of the
kernel sources it was measured on (comments and blank space ignored), and a change to those
sources without re-measuring fails here.  Re-measure with
`python -m mersenneprime_python_cpp_gmp_benchmark_amd.autotune --schedules` (large engine) or
`... --batch` (batched engine) on the GPU."""
import json
import os

import pytest

from mersenneprime_python_cpp_gmp_benchmark_amd import autotune


@pytest.mark.parametrize("path,kind", [(autotune.COSTS_PATH, "large"), (autotune.SCHED_JSON, "large"),
                                       (autotune.BATCH_JSON, "batch")],
                         ids=["length_costs", "large_schedules", "batch_costs"])
def test_cost_table_matches_kernel_sources(path, kind):
    with open(path) as f:
        d = json.load(f)
    assert d.get("kernel_hash") == autotune.kernel_source_hash(kind), (
        f"{os.path.basename(path)} was measured on other {kind}-engine kernels: re-run the autotuner")


def test_schedule_defaults_name_listed_schedules():
    """Every compiled-in schedule choice is a schedule of its length, and the fastest one measured."""
    with open(autotune.SCHED_JSON) as f:
        meas = json.load(f)["us_per_squaring"]
    rows = [l for l in open(autotune.SCHED_INC) if l.startswith("LLD(")]
    assert rows
    for line in rows:
        tr, n, w, h, cw = (int(v) for v in line[4:line.index(")")].split(","))
        key = f"{'fp64' if tr == 0 else 'ntt'}:{n}"
        times = meas[key]
        assert f"{w}x{h}/{cw}" in times
        assert times[f"{w}x{h}/{cw}"] == min(times.values())


def test_length_costs_cover_the_learned_candidates():
    costs = autotune.load_costs()
    loaded = autotune.load_costs(loaded=True)
    assert set(costs) == set(loaded)
    assert (1 << 16) in costs and (1 << 20) in costs and 917504 in costs
    assert all(v > 0 for v in costs.values()) and all(v > 0 for v in loaded.values())
