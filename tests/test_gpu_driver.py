"""The range driver on the device engines: every row's verdict and Res64 must equal the GMP
fixture, for each engine and for one call or many calls per list."""
import json

import pytest

from conftest import load_range_fixture
from mersenneprime_python_cpp_gmp_benchmark_amd import driver

pytestmark = pytest.mark.gpu


def _rows(path):
    return {r["p"]: (r["mersenne"], r["res64"]) for r in json.loads(path.read_text())["rows"]}


def _want(limit):
    return {p: (bool(v), f"{r:016x}") for p, v, r in load_range_fixture("c1_gmp_p10001.csv") if p <= limit}


@pytest.mark.parametrize("engine", ["fft", "gmp"])
def test_range_in_one_call_matches_gmp(tmp_path, engine):
    out = tmp_path / f"{engine}.json"
    assert driver.main(["-r", "10001", "--engine", engine, "--json", str(out), "--progress", "0"]) == 0
    assert _rows(out) == _want(10001)


def test_cuda_engine_in_chunks_matches_gmp(tmp_path):
    out = tmp_path / "cuda.json"
    assert driver.main(["-r", "5000", "--engine", "cuda", "--cuda-batch", "200", "--json", str(out),
                        "--progress", "0"]) == 0
    assert _rows(out) == _want(5000)


def test_list_skipping_known_exponents_exits_2(tmp_path):
    want = _want(10001)
    ps = [2, 3, 61, 89, 127, 521, 523, 1279, 2203, 4253, 4423, 7919, 9689, 9973]
    out = tmp_path / "list.json"
    rc = driver.main(["-l", *map(str, ps), "--json", str(out), "--progress", "0"])
    assert rc == 2      # 607, 2281, ... lie inside the list's span but are not in it
    assert _rows(out) == {p: want[p] for p in ps}


def test_verify_one_engine_against_another(tmp_path):
    ref = tmp_path / "ref.json"
    assert driver.main(["-r", "3000", "--engine", "cuda", "--json", str(ref), "--progress", "0"]) == 0
    assert driver.main(["-r", "3000", "--verify", str(ref), "--progress", "0"]) == 0


def test_loop_on_the_device(tmp_path):
    assert driver.main(["-r", "2000", "--loop", "2", "--progress", "0"]) == 0


def test_exponents_above_the_batch_limit(tmp_path):
    """The four primes just above 21 x 8192 exceed the batched FP64 lengths; the same call runs
    them as concurrent large-engine jobs."""
    want = {p: (bool(v), f"{r:016x}") for p, v, r in load_range_fixture("above_batch_gmp.csv")}
    out = tmp_path / "big.json"
    assert driver.main(["-l", *map(str, sorted(want)), "--json", str(out), "--progress", "0"]) == 0
    assert _rows(out) == want
