"""Every schedule of the large engine -- the H x W split of the transform and the column-pass
block width -- at every large-engine length of both transforms, through the C ABI: the Res64
after 201 squarings (two graph blocks, so the carry buffer between blocks is crossed) equals
GMP's, from the committed fixture tests/golden/sched_gmp.csv (one exponent per length: the
largest prime <= 16.5 N).  The schedule in force is restored after each length."""
import csv
import os

import pytest

from conftest import GOLDEN
from mersenneprime_python_cpp_gmp_benchmark_amd import llhip

pytestmark = pytest.mark.gpu

K = 201


def _fixture() -> dict:
    path = os.path.join(GOLDEN, "sched_gmp.csv")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(r["p"]): (int(r["iters"]), int(r["res64"], 16)) for r in csv.DictReader(f)}


FX = _fixture()


def _exponent(n: int) -> int:
    hi = int(16.5 * n)
    got = [p for p in FX if hi - 20000 <= p <= hi]
    return max(got) if got else 0


def _cases():
    out = []
    for tr in (llhip.FP64, llhip.NTT):
        for n in llhip.supported_lengths(tr):
            if llhip.large_schedules(n, tr):
                out.append((tr, n))
    return out


CASES = _cases()


@pytest.mark.parametrize("tr,n", CASES, ids=[f"{'ntt' if t else 'fp64'}-N{n}" for t, n in CASES])
def test_every_schedule_matches_gmp(tr, n):
    p = _exponent(n)
    if not p:
        pytest.skip("sched_gmp.csv has no exponent for this length")
    k, want = FX[p]
    assert k == K
    scheds = llhip.large_schedules(n, tr)
    assert scheds, n
    try:
        for s, (w, h, cw) in enumerate(scheds):
            llhip.set_large_schedule(n, s, tr)
            assert llhip.get_large_schedule(n, tr) == s
            r = llhip.lucas_lehmer_n(p, K, tr, n, no_rollback=True)
            assert r["status"] == llhip.ST_OK and r["n_used"] == n, (s, w, h, cw, r)
            assert r["res64"] == want, (s, w, h, cw, f"{r['res64']:016x}", f"{want:016x}")
            if tr == llhip.FP64:
                assert r["maxerr"] < 0.4, (s, r)
            print(f"{'ntt' if tr else 'fp64'} N={n} p={p} W={w} H={h} CW={cw}: "
                  f"{r['seconds'] / K * 1e6:.2f} us/squaring", flush=True)
    finally:
        llhip.set_large_schedule(n, -1, tr)

