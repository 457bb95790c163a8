"""Pins the CPU oracle of the learned-reluctivity MLP (oracle/mlp_oracle.py)."""
import numpy as np
import torch

from oracle import mlp_oracle as MO


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.0078125, 1.01171875]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(MO.bf16(x), ref)


def test_fp64_gradient_central_differences():
    """Backprop of the unrounded network against central differences."""
    rng = np.random.default_rng(2)
    p = MO.init_params(3) + rng.normal(size=MO.NPARAM) * 0.05
    X = rng.normal(size=(37, MO.N_IN))
    t = rng.normal(size=37)
    _, g = MO.loss_grad(p, X, t, bf=False)
    idx = np.concatenate([rng.choice(MO.NPARAM, 40, replace=False), [MO.NPARAM - 1, 0, 1024, 9408]])
    h = 1e-6
    for i in idx:
        e = np.zeros(MO.NPARAM)
        e[i] = h
        fd = (MO.loss_grad(p + e, X, t, bf=False)[0] - MO.loss_grad(p - e, X, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), (i, fd, g[i])


def test_bf16_network_close_to_fp64():
    """The BF16 rounding points move outputs and gradients by O(2^-8) only."""
    rng = np.random.default_rng(4)
    p = MO.init_params(5)
    X = rng.normal(size=(200, MO.N_IN))
    t = rng.normal(size=200)
    y64, _ = MO.forward(p, X, bf=False)
    yb, _ = MO.forward(p, X, bf=True)
    assert np.abs(yb - y64).max() < 3e-2 * np.abs(y64).max()
    _, g64 = MO.loss_grad(p, X, t, bf=False)
    _, gb = MO.loss_grad(p, X, t, bf=True)
    assert np.abs(gb - g64).max() < 3e-2 * np.abs(g64).max()


def test_tanh_rational_against_tanh():
    """The kernels' rational tanh (mirrored by the oracle) against numpy's:
    within 2.7e-7 everywhere (the clamp sets the largest error), odd, and
    1.3e-7 relative near zero; the BF16 rounding after it is 2^-9 relative."""
    x = np.concatenate([np.linspace(-12, 12, 200001), [0.0, 1e-30, -1e-20, 7.9, 8.5]])
    t = MO.tanh_rational(x)
    assert np.abs(t - np.tanh(x)).max() <= 2.7e-7
    assert np.array_equal(MO.tanh_rational(-x), -t)
    small = np.abs(x) < 1.0
    rel = np.abs(t[small] - np.tanh(x[small])) / np.maximum(np.abs(np.tanh(x[small])), 1e-300)
    assert rel.max() <= 1.3e-7
    assert np.all(np.abs(t) < 1.0)
