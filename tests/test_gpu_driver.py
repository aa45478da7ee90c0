"""The range driver (counterpart of PerfNumMultiCLL.py) on the device engines: every row's
verdict and Res64 must equal the GMP fixture, through each engine and dispatch form."""
import json

import pytest

from conftest import load_range_fixture
from mersenneprime_python_cpp_gmp_benchmark_amd import driver

pytestmark = pytest.mark.gpu


def _rows(path):
    return {r["p"]: (r["mersenne"], r["res64"]) for r in json.loads(path.read_text())["rows"]}


def _want(limit):
    return {p: (bool(v), f"{r:016x}") for p, v, r in load_range_fixture("c1_gmp_p10001.csv") if p <= limit}


def test_hip_engine_range_matches_gmp(tmp_path):
    out = tmp_path / "hip.json"
    assert driver.main(["-r", "10001", "--json", str(out), "--progress", "0"]) == 0
    assert _rows(out) == _want(10001)


def test_cuda_engine_in_chunks_matches_gmp(tmp_path):
    out = tmp_path / "cuda.json"
    assert driver.main(["-r", "5000", "--engine", "cuda", "--batch", "200", "--json", str(out), "--progress", "0"]) == 0
    assert _rows(out) == _want(5000)


@pytest.mark.parametrize("engine", ["fft", "gmp"])
def test_per_exponent_engines_match_gmp(tmp_path, engine):
    want = _want(10001)
    ps = [2, 3, 61, 89, 127, 521, 523, 1279, 2203, 4253, 4423, 7919, 9689, 9973]
    out = tmp_path / f"{engine}.json"
    rc = driver.main(["-l", *map(str, ps), "--engine", engine, "-t", "4", "--json", str(out), "--progress", "0"])
    assert rc == 2      # the list skips known exponents inside its span (607, 2281, ...), as the reference reports
    assert _rows(out) == {p: want[p] for p in ps}


def test_verify_against_another_engine(tmp_path):
    ref = tmp_path / "ref.json"
    assert driver.main(["-r", "3000", "--engine", "cuda", "--json", str(ref), "--progress", "0"]) == 0
    assert driver.main(["-r", "3000", "--verify", str(ref), "--progress", "0"]) == 0


def test_exponents_above_the_batch_limit(tmp_path):
    """The four primes just above 21 x 8192 exceed the batched FP64 lengths; the same batch call
    runs them as concurrent large-engine jobs."""
    want = {p: (bool(v), f"{r:016x}") for p, v, r in load_range_fixture("above_batch_gmp.csv")}
    out = tmp_path / "big.json"
    assert driver.main(["-l", *map(str, sorted(want)), "--json", str(out), "--progress", "0"]) == 0
    assert _rows(out) == want
