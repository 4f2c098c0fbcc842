"""Pins of the network oracle (oracle/mlp_oracle.c) by mathematics alone: central finite
differences of the FP64 network's loss against its analytic gradient, and the BF16-rounded
forward staying within bfloat16 precision of the FP64 forward."""
import numpy as np

import pyoracle as po


def make(seed=0, B=32):
    rng = np.random.default_rng(seed)
    n = po.mlp_nparam()
    assert n == 135 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4
    p = np.zeros(n, np.float32)
    o = 0
    for fan_in, rows, cols in ((135, 135, 64), (64, 64, 64), (64, 64, 4)):
        a = np.sqrt(6.0 / fan_in)
        p[o:o + rows * cols] = rng.uniform(-a, a, rows * cols)
        o += rows * cols
        p[o:o + cols] = rng.uniform(-0.1, 0.1, cols)
        o += cols
    X = rng.standard_normal((B, 135)).astype(np.float32)
    T = rng.standard_normal((B, 4)).astype(np.float32)
    return p, X, T


def test_fp64_backward_matches_finite_differences():
    p, X, T = make(1)
    g, loss = po.mlp_grad(p, X, T, bf16=False)
    rng = np.random.default_rng(2)
    # check a spread of parameters in every block (fp32 parameters -> step must stay representable)
    idx = np.concatenate([rng.choice(8640, 12, replace=False), 8640 + rng.choice(64, 4, replace=False),
                          8704 + rng.choice(4096, 8, replace=False), 12800 + rng.choice(64, 4, replace=False),
                          12864 + rng.choice(256, 8, replace=False), 13120 + np.arange(4)])
    checked = 0
    for i in idx:
        h = 1e-5 * max(1.0, abs(float(p[i])))
        p0 = p.astype(np.float64)
        pp, pm = p0.copy(), p0.copy()
        pp[i] += h
        pm[i] -= h
        pp32, pm32 = pp.astype(np.float32), pm.astype(np.float32)
        _, lp = po.mlp_grad(pp32, X, T, bf16=False)
        _, lm = po.mlp_grad(pm32, X, T, bf16=False)
        hp = float(pp32[i]) - float(p[i])
        hm = float(p[i]) - float(pm32[i])
        right, left = (lp - loss) / hp, (loss - lm) / hm
        if abs(right - left) > 1e-3 * max(1.0, abs(g[i])):
            continue  # a ReLU kink lies within +-h: one-sided slopes differ
        fd = (lp - lm) / (hp + hm)
        checked += 1
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])
    assert checked >= 0.8 * len(idx)


def test_bf16_forward_close_to_fp64():
    p, X, T = make(3, B=64)
    y64 = po.mlp_forward(p, X, bf16=False)
    y16 = po.mlp_forward(p, X, bf16=True)
    assert np.abs(y16 - y64).max() <= 0.03 * np.abs(y64).max()
    assert not np.array_equal(y16, y64)
