"""The reference's own test file (tests/test_lucaslehmer.py), run against the drop-in
`lucaslehmer` module whose engines run on the GPU; plus the lucaslehmer_hip batch API."""
import pytest

from conftest import KNOWN_4423

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ll():
    import lucaslehmer
    return lucaslehmer


def test_known_mersenne_exponents(ll):
    assert [p for p in ll.primes(2, 4423) if ll.lucas_lehmer(p)] == KNOWN_4423


def test_non_mersenne(ll):
    assert not ll.lucas_lehmer(11) and not ll.lucas_lehmer(23) and not ll.lucas_lehmer(4)


def test_compat_class(ll):
    c = ll.LucasLehmer()
    out = []
    assert c.sa_lucaslehmer(31, 4) is True and c.sa_getListOfPrimes(out, 2, 10) is True and out == [2, 3, 5, 7]


def test_split_matches_plain(ll):
    for d in (1, 2):
        assert [p for p in ll.primes(2, 4423) if ll.lucas_lehmer_split(p, d)] == KNOWN_4423
        assert not ll.lucas_lehmer_split(4425, d)


def test_fft_engine_matches_known_and_exact(ll):
    assert [p for p in ll.primes(2, 4423) if ll.lucas_lehmer_fft(p)] == KNOWN_4423
    assert all(ll.lucas_lehmer_fft(p) == ll.lucas_lehmer(p) for p in ll.primes(4425, 6000))
    n, bpw, err = ll.fft_info(9689)
    assert ll.lucas_lehmer_fft(9689) and err < 0.05 and bpw <= 18


def test_batch_module(ll):
    import lucaslehmer_hip as lc
    res = {p: r for p, r, s, res64 in lc.lucas_lehmer_batch(ll.primes(2, 4423)) if s >= 0}
    assert sorted(p for p, r in res.items() if r) == KNOWN_4423
    assert all(res[p] == ll.lucas_lehmer(p) for p in list(res)[::50])
    name, cus, mem = lc.device_info()
    assert cus > 0 and mem > 0


def test_res64_agrees_across_engines(ll):
    from oracle import oracle as O
    for p in (4425, 9689, 20011):
        g = ll.lucas_lehmer_res(p)
        f = ll.lucas_lehmer_fft_res(p)
        s = ll.lucas_lehmer_split_res(p, 1)
        assert g == f == s
        assert (g[1] == 0) == g[0]
        assert g == O.ll_gmp(p)[:2]


def test_interim_residues(ll):
    from oracle import oracle as O
    for p, k in ((9689, 1000), (44497, 777), (86243, 12345)):
        want = O.ll_gmp(p, k)[1]
        assert ll.gmp_res64_after(p, k) == want
        r, err = ll.fft_res64_after(p, k)
        assert r == want and err < 0.4
