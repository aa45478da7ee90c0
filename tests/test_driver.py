"""Host logic of the range driver, on the CPU.

A stand-in engine answers from the committed GMP fixture of every prime p <= 10001, so the rows,
the summary, the JSON file, --verify, --loop and the exit codes are exercised without a device;
test_gpu_driver.py runs the same driver on the device engines.
"""
import json

import pytest

from conftest import KNOWN_4423, load_range_fixture
from mersenneprime_python_cpp_gmp_benchmark_amd import driver


class FixtureEngine:
    """Answers each call from the fixture; `flip` corrupts one exponent's verdict and Res64."""

    def __init__(self, flip=None):
        self.rows = {p: (bool(v), r) for p, v, r in load_range_fixture("c1_gmp_p10001.csv")}
        self.flip = flip
        self.calls = []

    def __call__(self, ps):
        self.calls.append(list(ps))
        out = []
        for p in ps:
            v, r = self.rows[p]
            if p == self.flip:
                v, r = not v, r ^ 1
            out.append((p, v, 0.001, r))
        return out


def test_count_primes(capsys):
    assert driver.main(["-n", "2", "10001"]) == 0
    assert "primes in [2, 10001]: 1229" in capsys.readouterr().out


def test_worker_bounds():
    assert driver.main(["-t", "0", "-p", "7"], FixtureEngine()) == 1
    assert driver.main(["-t", str(driver.MAX_WORKERS + 1), "-p", "7"], FixtureEngine()) == 1


def test_json_rows_equal_fixture(tmp_path, capsys):
    out = tmp_path / "a.json"
    eng = FixtureEngine()
    assert driver.main(["-r", "10001", "--json", str(out), "--progress", "0"], eng) == 0
    assert len(eng.calls) == 1 and len(eng.calls[0]) == 1229          # one call for the whole list
    d = json.loads(out.read_text())
    want = load_range_fixture("c1_gmp_p10001.csv")
    assert [r["p"] for r in d["rows"]] == [p for p, _, _ in want]
    assert all(r["res64"] == f"{w:016x}" and r["mersenne"] == bool(v) for r, (_, v, w) in zip(d["rows"], want))
    assert d["hits"] == [p for p in KNOWN_4423 + [9689, 9941] if p <= 10001]
    assert {"python", "gil", "gmp", "workers", "executor", "engine", "fft_backend", "split", "order",
            "wall_s", "hits", "rows"} <= set(d)
    text = capsys.readouterr().out
    assert "tested 1229 exponents" in text and " YES " in text and "| yes" in text


def test_cuda_engine_chunks_ascending():
    eng = FixtureEngine()
    assert driver.main(["-r", "3000", "--engine", "cuda", "--cuda-batch", "100", "--progress", "0"], eng) == 0
    flat = [p for c in eng.calls for p in c]
    assert all(len(c) <= 100 for c in eng.calls) and flat == sorted(flat) and len(flat) == 430


def test_verify_catches_one_mismatch(tmp_path):
    ref = tmp_path / "ref.json"
    assert driver.main(["-r", "3000", "--json", str(ref), "--progress", "0"], FixtureEngine()) == 0
    assert driver.main(["-r", "3000", "--verify", str(ref), "--progress", "0"], FixtureEngine()) == 0
    d = json.loads(ref.read_text())
    d["rows"][100]["res64"] = "0123456789abcdef"
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps(d))
    assert driver.main(["-r", "3000", "--verify", str(bad), "--progress", "0"], FixtureEngine()) == 3


def test_loop_reference_then_checks(capsys):
    eng = FixtureEngine()
    assert driver.main(["-r", "2000", "--loop", "3", "--progress", "0"], eng) == 0
    assert len(eng.calls) == 3
    err = capsys.readouterr().err
    assert "pass 1: makes the reference" in err and "pass 3: checked against pass 1" in err


class FlipOnSecondPass(FixtureEngine):
    def __call__(self, ps):
        self.flip = 1009 if self.calls else None
        return super().__call__(ps)


def test_loop_stops_at_first_mismatch():
    eng = FlipOnSecondPass()
    assert driver.main(["-r", "2000", "--loop", "5", "--progress", "0"], eng) == 3
    assert len(eng.calls) == 2


def test_missing_known_exponent_exits_2():
    assert driver.main(["-r", "5000", "--progress", "0"], FixtureEngine(flip=4423)) == 2


@pytest.mark.parametrize("size,want", [(0, [[2, 3, 5, 7]]), (3, [[2, 3, 5], [7]]), (1, [[2], [3], [5], [7]])])
def test_chunked(size, want):
    assert driver.chunked([2, 3, 5, 7], size) == want
