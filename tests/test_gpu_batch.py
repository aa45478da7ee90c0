"""Parity of the batched GPU engines with the GMP oracle, through the C ABI.

Bit-exact: every verdict and every Res64 must equal the committed GMP fixtures.
"""
import numpy as np
import pytest

from conftest import KNOWN_100001, load_range_fixture
from mersenneprime_python_cpp_gmp_benchmark_amd import llhip

pytestmark = pytest.mark.gpu


def _check_rows(rows, out):
    ps = np.array([r[0] for r in rows], dtype=np.uint64)
    want_v = np.array([r[1] for r in rows])
    want_r = np.array([r[2] for r in rows], dtype=np.uint64)
    ok = np.isin(out["status"], [llhip.ST_OK, llhip.ST_TRIVIAL])
    assert ok.all(), ps[~ok][:10]
    bad_r = np.nonzero(out["res64"] != want_r)[0]
    bad_v = np.nonzero((out["verdict"] != 0) != (want_v != 0))[0]
    assert len(bad_r) == 0, [(int(ps[i]), hex(int(out["res64"][i])), hex(int(want_r[i]))) for i in bad_r[:10]]
    assert len(bad_v) == 0, ps[bad_v][:10]


@pytest.mark.parametrize("transform", [llhip.NTT, llhip.FP64], ids=["ntt", "fp64"])
def test_c1_all_primes_to_10001_one_launch(transform):
    rows = load_range_fixture("c1_gmp_p10001.csv")
    out = llhip.batch([r[0] for r in rows], 0, transform)
    _check_rows(rows, out)
    if transform == llhip.FP64:
        assert float(out["maxerr"].max()) < 0.4


@pytest.mark.parametrize("transform", [llhip.NTT, llhip.FP64], ids=["ntt", "fp64"])
def test_c3_all_primes_to_100001_one_launch(transform):
    rows = load_range_fixture("c3_gmp_p100001.csv")
    out = llhip.batch([r[0] for r in rows], 0, transform)
    _check_rows(rows, out)
    hits = [int(p) for p, v in zip([r[0] for r in rows], out["verdict"]) if v]
    assert hits == KNOWN_100001


def test_partial_iterations_match_gmp():
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    ps = [int(p) for p in rng.choice(O.primes(65, 30000), 64, replace=False)]
    ks = [int(rng.integers(1, p - 2)) for p in ps]
    for tr in (llhip.NTT, llhip.FP64):
        for p, k in zip(ps[:16], ks[:16]):
            r = llhip.lucas_lehmer(p, k, tr)
            assert r["status"] == llhip.ST_OK
            assert r["res64"] == O.ll_gmp(p, k)[1], (tr, p, k)


def test_odd_composite_and_tiny_exponents():
    from oracle import oracle as O
    ps = [3, 5, 7, 9, 11, 15, 21, 23, 25, 29, 31, 33, 45, 49, 61, 63] + list(range(65, 1200, 2)) + [4425, 6001]
    for tr in (llhip.NTT, llhip.FP64):
        out = llhip.batch(ps, 0, tr)
        for i, p in enumerate(ps):
            v, r, _ = O.ll_gmp(p)
            assert (bool(out["verdict"][i]), int(out["res64"][i])) == (v, r), (tr, p)


def test_argument_rules():
    out = llhip.batch([0, 1, 2, 4, 100], 0, llhip.NTT)
    assert list(out["verdict"]) == [0, 0, 1, 0, 0]
    assert all(s == llhip.ST_TRIVIAL for s in out["status"])


def test_forced_rollback_stays_exact():
    """Too many bits per word makes the FP64 round-off exceed 0.4; the engine must roll back
    to longer lengths and still return GMP's Res64."""
    rows = load_range_fixture("c1_gmp_p10001.csv")
    sel = [r for r in rows if r[0] > 3000][::40]
    old = llhip.get_batch_bpw()
    try:
        llhip.set_batch_bpw(40.0)
        first = np.array([llhip.lib().llh_batch_length(r[0], llhip.FP64) for r in sel])
        out = llhip.batch([r[0] for r in sel], 0, llhip.FP64)
    finally:
        llhip.set_batch_bpw(old)
    _check_rows(sel, out)
    assert (out["n_used"] > first).any()          # some exponents were rolled back to 2N
    assert (out["n_used"] >= first).all()


def _fp64_ladder():
    lens = llhip.supported_lengths(llhip.FP64)
    return [n for n in lens if n <= 8192]


def test_every_batch_length_matches_gmp():
    """Every FP64 batch length (powers of two and 3, 5, 7 times a power of two), forced onto the
    three largest primes within the batch cap of bits per digit, in one batch per length: Res64 after 1500
    squarings (and, for the shorter lengths, after the full test) must equal GMP's."""
    from oracle import oracle as O
    lad = _fp64_ladder()
    assert {384, 640, 896, 1536, 2560, 3584, 5120, 6144, 7168} <= set(lad)
    for n in lad:
        if n < 64:
            continue
        ps = [p for p in O.primes(max(65, int(n * 14)), int(n * llhip.get_batch_bpw())) if (n + 2) * p < (1 << 32)][-3:]
        out = llhip.batch_n(ps, 1500, llhip.FP64, n, no_rollback=True)
        assert (out["status"] == llhip.ST_OK).all(), (n, ps, list(out["status"]))
        assert (out["n_used"] == n).all(), n
        for i, p in enumerate(ps):
            assert int(out["res64"][i]) == O.ll_gmp(p, 1500)[1], (p, n)
        if ps[-1] < 12000:
            full = llhip.batch_n(ps, 0, llhip.FP64, n, no_rollback=True)
            for i, p in enumerate(ps):
                v, r, _ = O.ll_gmp(p)
                assert (bool(full["verdict"][i]), int(full["res64"][i])) == (v, r), (p, n)


def test_length_choice_is_cheapest_within_cap():
    """The engine's length for p is, among the lengths within the bits-per-digit cap (up to
    twice the shortest), the one with the lowest measured cost."""
    L = llhip.lib()
    cap = llhip.get_batch_bpw()
    lad = _fp64_ladder()
    for p in (9689, 11213, 21701, 44497, 60013, 86243, 99991):
        n = L.llh_batch_length(p, llhip.FP64)
        ok = [m for m in lad if p / m <= cap]
        cands = [m for m in ok if m <= 2 * ok[0]]
        costs = {m: llhip.get_batch_cost(m, llhip.FP64) for m in cands}
        assert n in cands, (p, n)
        assert costs[n] == min(costs.values()), (p, n, costs)


def test_odd_lengths_forced_match_gmp():
    """One exponent forced onto every batch length between 6 and 20 bits per digit gives one
    Res64 (below about 4 bits per digit the two-round parallel carry no longer balances the
    digits and the round-off check rolls the exponent back)."""
    from oracle import oracle as O
    for p, k in ((24989, 2000), (99991, 1000)):
        want = O.ll_gmp(p, k)[1]
        for n in _fp64_ladder():
            if not 6.0 <= p / n <= 20.0:
                continue
            r = llhip.lucas_lehmer_n(p, k, llhip.FP64, n)
            assert r["status"] == llhip.ST_OK and r["n_used"] == n, (p, n, r)
            assert r["res64"] == want, (p, n)


def test_carry_rounds_both_sides_of_threshold():
    """The FP64 carry takes one hand-off round from log2 N + 3 bits per digit and two below it;
    exponents just either side of that boundary, forced onto the length, must match GMP."""
    from oracle import oracle as O
    for n in (512, 4096, 5120):
        thr = ((n - 1).bit_length() + 3) * n            # ceil(log2 N) + 3 bits per digit
        below = [p for p in O.primes(thr - 400, thr - 1)][-2:]
        above = [p for p in O.primes(thr, thr + 400)][:2]
        ps = below + above
        out = llhip.batch_n(ps, 2000, llhip.FP64, n, no_rollback=True)
        assert (out["status"] == llhip.ST_OK).all(), (n, ps, list(out["status"]))
        assert (out["n_used"] == n).all(), n
        for i, p in enumerate(ps):
            assert int(out["res64"][i]) == O.ll_gmp(p, 2000)[1], (p, n)
        assert float(np.max(out["maxerr"])) < 0.1, (n, list(out["maxerr"]))


def _c3_subset():
    rows = load_range_fixture("c3_gmp_p100001.csv")
    return [r for r in rows if r[0] > 20000][::6]


def test_rollbacks_resume_from_checkpoints():
    """At 22 bits per digit the batch lengths' round-off crosses 0.4 mid-run for some exponents.
    Each must roll back to the next length from its last good checkpoint (at most one check
    interval, 64 squarings, redone), not from s_0, and every Res64 must equal GMP's."""
    sel = _c3_subset()
    ps = [r[0] for r in sel]
    total = sum(p - 2 for p in ps)
    old = llhip.get_batch_bpw()
    assert not llhip.get_batch_checkpoints()
    try:
        llhip.set_batch_bpw(22.0)
        llhip.set_batch_checkpoints(True)
        out = llhip.batch(ps, 0, llhip.FP64)
        st = llhip.last_run_stats()
        llhip.set_batch_checkpoints(False)
        out0 = llhip.batch(ps, 0, llhip.FP64)
        st0 = llhip.last_run_stats()
    finally:
        llhip.set_batch_checkpoints(False)
        llhip.set_batch_bpw(old)
    _check_rows(sel, out)
    _check_rows(sel, out0)
    assert st["rollbacks"] > 0 and st["resumed"] == st["rollbacks"], st
    assert total <= st["squarings"] <= total + 64 * st["rollbacks"], (st, total)
    assert st0["rollbacks"] > 0 and st0["resumed"] == 0 and st0["squarings"] > st["squarings"]


def test_every_batch_schedule_matches_gmp():
    """Every schedule (radix order) of every batched length that has more than one, on both
    transforms where it runs: full tests of twelve primes at that length (p / N from 16 bits to
    the cap, or to the NTT's exactness bound) give GMP's verdicts and Res64 (the C3 fixture).
    The schedule in force is restored after each length."""
    fx = {p: (v, r) for p, v, r in load_range_fixture("c3_gmp_p100001.csv")}
    primes = sorted(fx)
    L = llhip.lib()
    checked = 0
    for tr in (llhip.FP64, llhip.NTT):
        for n in llhip.supported_lengths(tr):
            scheds = llhip.batch_schedules(n, tr)
            if len(scheds) < 2:
                continue
            if tr == llhip.FP64:
                lo, hi = 16 * n, int(llhip.get_batch_bpw() * n)
            else:
                bits = (63 - (n.bit_length() - 1)) // 2
                lo, hi = int(0.7 * bits * n), bits * n - n // 2
            cand = [p for p in primes if lo <= p <= min(hi, 100001)]
            ps = cand[:: max(1, len(cand) // 12)][:12]
            assert len(ps) >= 6, (tr, n)
            try:
                for s, rad in enumerate(scheds):
                    llhip.set_batch_schedule(n, s, tr)
                    out = llhip.batch_n(ps, 0, tr, n, no_rollback=True)
                    for i, p in enumerate(ps):
                        assert int(out["status"][i]) == llhip.ST_OK and int(out["n_used"][i]) == n, (tr, n, rad, p)
                        assert (int(out["verdict"][i] != 0), int(out["res64"][i])) == fx[p], (tr, n, rad, p)
                    checked += 1
            finally:
                llhip.set_batch_schedule(n, -1, tr)
    assert checked >= 10
