"""Learned FFT-length choice for the FP64 engine.

A small MLP (3 -> 64 -> 64 -> 1, FP32) predicts log10 of the worst round-off an FP64
IBDWT run will see from (log2 p, log2 N, bits per digit).  It is trained on the GPU by the
library's own kernels (llh_mlp_train) on round-off the engine itself measured
(roundoff_data.py), and picks the smallest supported length whose predicted round-off
stays under a safety bound.  The per-iteration round-off check of the engine (above 0.4:
rollback to the next length) remains the fallback and the judge of every choice.

The reference instead takes the smallest {1,3,5} * 2^k length with at most 18 bits per
digit (lucaslehmer.cpp:244-254); that rule is `conservative_length`.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "roundoff_mlp.json")
DATA_PATH = os.path.join(HERE, "roundoff_train.csv")

# predicted worst round-off over the measured window must stay below this (the engine's
# rollback threshold is 0.4; the margin covers the slow growth of the maximum over a long run
# and the model's own error)
SAFE_ERR = 0.15
# the window the targets were measured over
WINDOW_ITERS = 2000
MIN_LARGE_N = 1 << 16


def features(p, n) -> np.ndarray:
    """Normalised (log2 p, log2 N, bits per digit) rows."""
    p = np.atleast_1d(np.asarray(p, np.float64))
    n = np.atleast_1d(np.asarray(n, np.float64))
    return np.stack([(np.log2(p) - 22.5) / 1.5, (np.log2(n) - 19.5) / 1.5, (p / n - 17.0) / 2.0], 1).astype(np.float32)


def target_of(maxerr) -> np.ndarray:
    return ((np.log10(np.maximum(np.asarray(maxerr, np.float64), 1e-6)) + 2.0) / 1.0).astype(np.float32)


def err_of(target) -> np.ndarray:
    return 10.0 ** (np.asarray(target, np.float64) * 1.0 - 2.0)


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, in the kernels' layout."""
    rng = np.random.default_rng(seed)
    H = 64
    def glorot(fan_in, fan_out, shape):
        a = math.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-a, a, shape)
    parts = [glorot(3, H, (H, 3)).ravel(), np.zeros(H), glorot(H, H, (H, H)).ravel(), np.zeros(H),
             glorot(H, 1, (H,)), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def conservative_length(p: int, bpw: float = 18.0) -> int:
    """The reference's rule (lucaslehmer.cpp:244-254)."""
    best = 0
    for m in (1, 3, 5):
        n = m * 16
        while n < (1 << 40):
            if p / n <= bpw:
                best = n if not best or n < best else best
                break
            n <<= 1
    return best


def load_data(path: str = DATA_PATH):
    rows = [l.strip().split(",") for l in open(path)][1:]
    p = np.array([int(r[0]) for r in rows], np.float64)
    n = np.array([int(r[1]) for r in rows], np.float64)
    e = np.array([float(r[2]) for r in rows], np.float64)
    return p, n, e


def train(epochs: int = 4000, lr: float = 3e-3, seed: int = 0, path: str = DATA_PATH) -> dict:
    """Train on the GPU (llh_mlp_train) and return the model dict (also written to MODEL_PATH)."""
    from . import llhip
    p, n, e = load_data(path)
    X = features(p, n)
    t = target_of(e)
    params, hist = llhip.mlp_train(init_params(seed), X, t, epochs, lr)
    pred = llhip.mlp_predict(params, X)
    model = {"params": [float(v) for v in params], "epochs": epochs, "lr": lr, "seed": seed,
             "samples": int(len(t)), "final_loss": float(hist[-1]), "first_loss": float(hist[0]),
             "rmse_log10": float(np.sqrt(np.mean((pred - t) ** 2))), "safe_err": SAFE_ERR,
             "window_iters": WINDOW_ITERS}
    with open(MODEL_PATH, "w") as f:
        json.dump(model, f)
    return model


def load_model(path: str = MODEL_PATH) -> dict:
    with open(path) as f:
        return json.load(f)


def predict_err(p, n, model: dict | None = None) -> np.ndarray:
    """Predicted worst round-off (GPU inference kernel)."""
    from . import llhip
    model = model or load_model()
    return err_of(llhip.mlp_predict(np.array(model["params"], np.float32), features(p, n)))


def candidate_lengths() -> list[int]:
    from . import llhip
    return [n for n in llhip.supported_lengths(llhip.FP64) if n >= MIN_LARGE_N]


MIN_GAIN = 0.03     # leave the conservative length only for a measured gain of 3 % or more


def choose_length(p: int, model: dict | None = None, safe_err: float = SAFE_ERR, costs: dict | None = None,
                  min_gain: float = MIN_GAIN, loaded: bool = False) -> int:
    """The cheapest supported FP64 length whose predicted round-off is below safe_err (cost
    per squaring from the autotuner's table, shortest length when there is no table), never
    beyond the bits-per-digit range the model was trained on; the conservative length if
    nothing qualifies.  The conservative length is always a candidate when it is safe, and it
    is kept unless the cheapest candidate is at least min_gain cheaper (the table's costs of
    neighbouring lengths can differ by less than their run-to-run noise).  loaded: price the
    lengths with several exponents in flight (one call over many exponents)."""
    model = model or load_model()
    cons = conservative_length(p)
    cands = sorted(set(n for n in candidate_lengths() if 14.0 <= p / n <= 21.0) | {cons})
    if costs is None:
        try:
            from .autotune import load_costs
            costs = load_costs(loaded=loaded)
        except (OSError, ValueError, KeyError):
            costs = {}
    err = predict_err(np.full(len(cands), p), np.array(cands), model)
    safe = [int(n) for n, e in zip(cands, err) if e <= safe_err or n == cons and p / n <= 18.0]
    if not safe:
        return cons
    if costs and all(n in costs for n in safe):
        best = min(safe, key=lambda n: (costs[n], n))
        if cons in safe and cons in costs and costs[best] > (1.0 - min_gain) * costs[cons]:
            return cons
        return best
    return min(safe)
