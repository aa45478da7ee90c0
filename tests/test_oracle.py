"""Pins the CPU oracle (oracle/ll_oracle.c) -- CPU only.

The oracle is the checker for every GPU parity test, so it is itself checked against:
the reference's known-answer tests (tests/test_lucaslehmer.py), known_mp_list.txt,
Python big-int brute force, the closed form of the unreduced LL sequence, textbook
Goldilocks arithmetic, and agreement of its three independent engines (GMP, NTT, FP64).
"""
import random

import pytest

from conftest import KNOWN_4423, KNOWN_100001, load_range_fixture
from oracle import oracle as O

Q = 2**64 - 2**32 + 1


def test_sieve_matches_reference_tests():
    # tests/test_lucaslehmer.py:5-9
    assert len(O.primes(2, 10001)) == 1229
    assert O.primes(0, 1) == [] and O.primes(2, 2) == [2] and O.primes(10, 20) == [11, 13, 17, 19]
    assert len(O.primes(2, 100001)) == 9592


def test_known_mersenne_exponents_gmp():
    # tests/test_lucaslehmer.py:11-13
    assert [p for p in O.primes(2, 4423) if O.ll_gmp(p)[0]] == KNOWN_4423


def test_non_mersenne():
    # tests/test_lucaslehmer.py:15-16
    assert not O.ll_gmp(11)[0] and not O.ll_gmp(23)[0] and not O.ll_gmp(4)[0]


def test_gmp_equals_python_bigint_bruteforce():
    for p in O.primes(3, 700):
        v, r, deg = O.ll_gmp(p)
        assert (v, r) == O.py_ll(p), p
        assert not deg
    rng = random.Random(5)
    for _ in range(40):
        p = rng.choice(O.primes(3, 3000))
        k = rng.randrange(1, p)
        assert O.ll_gmp(p, k)[:2] == O.py_ll(p, k)


def test_unreduced_closed_form():
    # while s_k < M_p no reduction happens: s_k = (2+sqrt3)^(2^k) + (2-sqrt3)^(2^k), i.e. the plain
    # integer recurrence; checked against the exact integer sequence
    p = 4423
    s = 4
    for k in range(1, 12):
        s = s * s - 2
        assert s < 2**p - 1
        assert O.ll_state(p, k) == s


def test_res64_zero_iff_prime():
    # tests/test_lucaslehmer.py:53
    for p in O.primes(3, 2500):
        v, r, _ = O.ll_gmp(p)
        assert (r == 0) == v


def test_three_engines_agree():
    # tests/test_lucaslehmer.py:49-53 (Res64 across engines) at p in {4425, 9689, 20011} and more
    for p in (521, 607, 1279, 4425, 9689):
        g = O.ll_gmp(p)[:2]
        nv, nr = O.ll_ntt(p, O.ntt_length(p))
        fv, fr, err = O.ll_fp64(p, O.fft_length(p), abort_err=0.35)
        assert (bool(nv), nr) == g == (bool(fv), fr), p
        assert err < 0.05


def test_fixture_c1_consistent():
    rows = load_range_fixture("c1_gmp_p10001.csv")
    assert [p for p, _, _ in rows] == O.primes(2, 10001)
    assert [p for p, v, _ in rows if v] == [p for p in KNOWN_100001 if p <= 10001]
    assert all((r == 0) == bool(v) for p, v, r in rows if p > 2)
    rng = random.Random(1)
    for p, v, r in rng.sample(rows, 25):
        assert O.ll_gmp(p)[:2] == (bool(v), r)


def test_fixture_c3_consistent():
    rows = load_range_fixture("c3_gmp_p100001.csv")
    assert [p for p, _, _ in rows] == O.primes(2, 100001)
    assert [p for p, v, _ in rows if v] == KNOWN_100001
    rng = random.Random(2)
    small = [row for row in rows if row[0] < 20000]
    for p, v, r in rng.sample(small, 10):
        assert O.ll_gmp(p)[:2] == (bool(v), r)


def test_goldilocks_field():
    rng = random.Random(3)
    for _ in range(200):
        a, b = rng.randrange(Q), rng.randrange(Q)
        assert O.gl_mul(a, b) == a * b % Q
    # 2 has order 192; the N-th root of 2 exists for N | 2^26 (cuda/ll_ntt.cu:8-11, :189-192)
    assert pow(2, 192, Q) == 1 and pow(2, 96, Q) == Q - 1
    for n in (16, 1024, 1 << 20):
        w = O.gl_root2(n)
        assert pow(w, n, Q) == 2
    assert O.gl_root2(3 << 10) == 0     # no 3*2^k root of 2 in this field


def test_digit_layout_invariants():
    for p, n in ((1048573, 65536), (16777213, 1 << 20), (100003, 8192), (521, 32)):
        bits, we = O.digit_layout(p, n) if n <= 8192 else (None, None)
        if bits is None:
            continue
        assert sum(bits) == p
        lo, hi = p // n, -(-p // n)
        assert set(bits) <= {lo, hi}
        assert all(0 <= e < n for e in we)


BATCH_M = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 192, 384, 768, 1536, 3072,
           320, 640, 1280, 2560, 448, 896, 1792, 3584)


def test_digit_bits_closed_form():
    """The batch engine's digit walk: digit d of weight exponent e = (-d p) mod N has
    p div N + [e < p mod N] bits (p prime and coprime to N), checked against the ceiling
    definition (lucaslehmer.cpp:262-266) and the oracle's layout at every batch length."""
    rng = random.Random(11)
    cands = [q for q in O.primes(65, 180000)]
    for m in BATCH_M:
        for n in (2 * m, m):
            ps = [q for q in cands if n < q < 21 * n]
            for p in rng.sample(ps, min(6, len(ps))) + [n + 1]:
                bits, we = O.digit_layout(p, n)
                P, F = p % n, p // n
                for d in range(n):
                    e = (-d * p) % n
                    want = -(-(d + 1) * p // n) - (-(-d * p // n))
                    assert F + (e < P) == want == bits[d], (n, p, d)
                    assert we[d] == e, (n, p, d)


def test_split_gmp_engine():
    """The split-Karatsuba engine (lucaslehmer.cpp:83-179) finds exactly the known Mersenne
    exponents up to 4423 at depths 1 and 2 (tests/test_lucaslehmer.py:26-29) and gives the plain
    GMP Res64 on composites above the 4096-bit split threshold, also after a partial run."""
    ps = O.primes(2, 4423)
    for d in (1, 2):
        assert [p for p in ps if p > 4000 and O.ll_gmp_split(p, 0, d)[0]] == [4253, 4423]
        assert not O.ll_gmp_split(4425, 0, d)[0]
        for p in (9689, 9973, 21701):
            assert O.ll_gmp_split(p, 0, d) == O.ll_gmp(p)[:2], (p, d)
        assert O.ll_gmp_split(44497, 300, d) == O.ll_gmp(44497, 300)[:2]
    assert O.bench_square(1 << 16, 5, 1) > 0.0


def test_size_rules():
    # lucaslehmer.cpp:244-254 at the configurations' exponents
    assert O.fft_length(1048573) == 65536
    assert O.fft_length(16777213) == 1 << 20
    assert O.fft_length(100001) == 6144
    assert O.fft_length(9689) == 640
    # cuda/ll_ntt.cu:172-180
    assert O.ntt_length(100001) == 4096
    assert O.ntt_length(64) == 64
    assert O.ntt_length(205000) == 0


def test_relayout_roundtrip():
    rng = random.Random(4)
    for p, n in ((4423, 256), (9689, 512), (521, 64)):
        x = rng.randrange(2**p - 1)
        d = O.limbs_to_digits(p, n, x)
        assert O.digits_res64(p, d)[1] == x & (2**64 - 1)


@pytest.mark.parametrize("p,k", [(4423, 100), (9689, 57)])
def test_interim_engines_agree(p, k):
    g = O.ll_gmp(p, k)[1]
    assert O.ll_ntt(p, O.ntt_length(p), k)[1] == g
    assert O.ll_fp64(p, O.fft_length(p), k)[1] == g
