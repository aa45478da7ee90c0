"""The two drop-in modules on the device, checked against the committed GMP fixture and the oracle.

`lucaslehmer` must answer every call of the reference module's surface (lucaslehmer.cpp:344-387)
with the GMP engine's verdict and Res64; `lucaslehmer_cuda` must serve the reference caller's
dispatch loop (PerfNumMultiCLL.py:140-157) unchanged: chunks by `ntt_length`, the seconds < 0
redo, and (p, is_prime, seconds, res64) rows.
"""
import threading

import pytest

from conftest import load_range_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ll():
    import lucaslehmer
    return lucaslehmer


@pytest.fixture(scope="module")
def c1():
    return {p: (bool(v), r) for p, v, r in load_range_fixture("c1_gmp_p10001.csv")}


def _sample(c1, step):
    ps = sorted(c1)
    return ps[1::step] + [p for p in ps if c1[p][0]]


def test_exact_calls_equal_gmp_fixture(ll, c1):
    for p in _sample(c1, 37):
        v, r = c1[p]
        assert ll.lucas_lehmer(p) is v, p
        assert ll.lucas_lehmer_res(p) == (v, r), p


def test_fft_calls_equal_gmp_fixture(ll, c1):
    for p in _sample(c1, 41):
        v, r = c1[p]
        assert ll.lucas_lehmer_fft(p) is v, p
        assert ll.lucas_lehmer_fft_res(p) == (v, r), p


def test_split_calls_equal_gmp_fixture(ll, c1):
    for depth in (1, 2):
        for p in _sample(c1, 97):
            assert ll.lucas_lehmer_split_res(p, depth) == c1[p], (depth, p)
            assert ll.lucas_lehmer_split(p, depth) is c1[p][0]


def test_interim_residues_equal_oracle(ll):
    from oracle import oracle as O
    for p, k in ((521, 1), (9689, 1000), (44497, 777), (86243, 12345), (132049, 3001)):
        want = O.ll_gmp(p, k)[1]
        assert ll.gmp_res64_after(p, k) == want, (p, k)
        r, err = ll.fft_res64_after(p, k)
        assert r == want and 0.0 <= err < 0.4, (p, k, err)


def test_fft_info_follows_the_bits_per_word_knob(ll):
    from oracle import oracle as O
    try:
        for bits, p in ((18.0, 9689), (16.0, 44497), (12.0, 20011)):
            ll.set_fft_bits_per_word(bits)
            n, bpw, err = ll.fft_info(p)
            assert n == O.fft_length(p, bits) and abs(bpw - p / n) < 1e-9 and bpw <= bits
            assert 0.0 < err < 0.4
    finally:
        ll.set_fft_bits_per_word(18.0)


def test_argument_rules_without_device_state(ll):
    assert ll.lucas_lehmer(0) is False and ll.lucas_lehmer(1) is False and ll.lucas_lehmer(2) is True
    assert ll.lucas_lehmer_res(2) == (True, 0) and ll.lucas_lehmer_res(1000) == (False, 0)
    assert ll.lucas_lehmer_fft_res(2) == (True, 0)
    with pytest.raises(TypeError):
        ll.lucas_lehmer(-5)


def test_concurrent_python_threads_get_their_own_results(ll, c1):
    ps = _sample(c1, 53)
    got = {}

    def work(chunk):
        for p in chunk:
            got[p] = ll.lucas_lehmer_res(p)
    ts = [threading.Thread(target=work, args=(ps[i::4],)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert got == {p: c1[p] for p in ps}


def test_reference_dispatch_loop_over_c1(ll, c1):
    """PerfNumMultiCLL.py:140-157, restated as the caller would run it against this module."""
    import lucaslehmer_cuda as lc
    name, cus, mem = lc.device_info()
    assert isinstance(name, str) and cus > 0 and mem > 0
    chunks = []
    for p in sorted(c1):
        n = lc.ntt_length(p) if p >= 64 else 0
        if chunks and chunks[-1][0] == n and len(chunks[-1][1]) < 300:
            chunks[-1][1].append(p)
        else:
            chunks.append((n, [p]))
    rows = {}
    for n, ps in chunks:
        res = {p: (r, dt, rs) for p, r, dt, rs in lc.lucas_lehmer_batch(ps)} if n else {}
        for p in ps:
            if p in res and res[p][1] >= 0:
                r, _dt, rs = res[p]
            else:
                r, rs = ll.lucas_lehmer_fft_res(p)
            rows[p] = (r, rs)
    assert rows == c1
    assert {n for n, _ in chunks} == {0, 64, 128, 256, 512}     # the reference rule's lengths <= 10001


def test_batch_rows_types_and_order(c1):
    import lucaslehmer_cuda as lc
    ps = [9973, 3, 127, 8191, 61, 4423, 9941]
    for transform in ("ntt", "fp64"):
        out = lc.lucas_lehmer_batch(ps, 0, transform)
        assert [row[0] for row in out] == ps
        for p, prime, secs, res in out:
            assert isinstance(prime, bool) and isinstance(secs, float) and isinstance(res, int)
            assert secs >= 0 and (prime, res) == c1[p], (transform, p)


def test_batch_beyond_the_reference_bound_equals_gmp():
    """p above 204,800 (the reference kernel's bound) and in (172,032, 884,736], where the FP64
    conservative length lies below the large engine's shortest; 50 squarings each."""
    import lucaslehmer_cuda as lc
    from oracle import oracle as O
    ps = [O.primes(n, n + 200)[0] for n in (200003, 500009, 786433, 884700)]
    want = {p: O.ll_gmp(p, 50)[1] for p in ps}
    for transform in ("fp64", "ntt"):
        out = lc.lucas_lehmer_batch(ps, 50, transform)
        assert [(p, res) for p, _, secs, res in out if secs >= 0] == list(want.items()), transform
