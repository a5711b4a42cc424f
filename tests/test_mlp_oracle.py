"""CPU: the FP64 MLP oracle's analytic backprop against central differences."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def _params(rng):
    shapes = M.layer_shapes()
    parts = []
    for (o, i) in shapes:
        parts.append(rng.uniform(-1, 1, o * i) * np.sqrt(6.0 / i))
        parts.append(rng.uniform(-0.1, 0.1, o))
    return np.concatenate(parts)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    freq = rng.normal(size=(64, 3))
    pos = rng.uniform(-1, 1, (12, 3))
    idx = rng.integers(0, 4, 12)
    tgt = rng.normal(size=12)
    loss, g = M.loss_and_grad(p, freq, pos, idx, tgt, bf=False)
    # a few coordinates in every layer (weights and biases)
    off = 0
    h = 1e-6
    for (o, i) in M.layer_shapes():
        for k in list(rng.integers(off, off + o * i, 4)) + list(rng.integers(off + o * i, off + o * i + o, 2)):
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            lp, _ = M.loss_and_grad(pp, freq, pos, idx, tgt, bf=False)
            lm, _ = M.loss_and_grad(pm, freq, pos, idx, tgt, bf=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
        off += o * i + o
    assert off == len(p)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8
