"""The MLP oracle (oracle/mlp_oracle.py) pinned by mathematics: its analytic
gradient against central finite differences of its own float64 loss, the
bfloat16 rounding against known bit patterns, and GELU against its closed form
at special points."""
import math

import numpy as np

import mlp_oracle as M


def _data(n=96, seed=0):
    rng = np.random.default_rng(seed)
    X = np.stack([rng.uniform(1e7, 3e7, n), rng.uniform(-4e7, -1e7, n), rng.uniform(0.3, 0.9, n),
                  rng.uniform(0.008, 0.016, n), np.full(n, 0.012), rng.uniform(-28, 1, n)], axis=1)
    y = rng.uniform(-28, 2, n)
    return X, y


def test_parameter_count():
    assert M.NTRAIN == 6 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 8833
    assert M.NPARAMS == 8833 + 14


def test_gradient_matches_finite_differences():
    X, y = _data()
    p = M.init_params(X, y, seed=3).astype(np.float64)
    loss, g = M.loss_and_grad(p, X, y, bf16_points=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(M.NTRAIN, 60, replace=False), [M.NTRAIN - 1, 0, 384, 8832 - 64]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_and_grad(pp, X, y, False)[0] - M.loss_and_grad(pm, X, y, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_bf16_rounding():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to even (1.0)
    assert M.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    # 1 + 3*2^-8 ties between 1+2^-7 and 1+2^-6: even is 1+2^-6
    assert M.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert M.bf16(np.array([1.0 + 2.0 ** -8 + 2.0 ** -20]))[0] == 1.0 + 2.0 ** -7
    x = np.array([3.0, -0.5, 1024.0, 0.0])
    assert np.array_equal(M.bf16(x), x)


def test_gelu_special_values():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([1.0]))[0] - 0.5 * (1 + math.erf(1 / math.sqrt(2)))) < 1e-15
    assert abs(M.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    z = np.linspace(-5, 5, 41)
    fd = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-8


def test_bf16_points_change_little():
    X, y = _data()
    p = M.init_params(X, y, seed=5)
    a = M.forward(p, X, True)
    b = M.forward(p, X, False)
    assert np.max(np.abs(a - b)) < 0.05 * max(1.0, np.max(np.abs(b)))
