"""Pins the learned-nprobe oracle (oracle/mlp_oracle.py) on the CPU:
bf16 rounding against torch's bfloat16 conversion, backward against finite
differences in float64, the covering-nprobe label against its definition."""
import numpy as np
import pytest

import mlp_oracle as M


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.0078125, 1.005859375, -2.5e-39, 3.4e38],
                                 np.float32)])
    ours = M.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)


@pytest.mark.parametrize("seed,tau", [(0, None), (1, None), (2, 0.9), (3, 0.25)])
def test_backward_matches_finite_differences(seed, tau):
    rng = np.random.default_rng(seed)
    F, H, B = 8, 16, 12
    p = M.init(F, H, seed)
    for k in ("b1", "b2", "b3"):
        p[k] = rng.standard_normal(p[k].shape) * 0.1
    X = rng.standard_normal((B, F))
    y = rng.standard_normal(B)
    z, cache = M.forward(p, X, bf=False)
    _, g = M.loss_grad(z, y, bf=False, tau=tau)
    grads = M.backward(cache, g, bf=False)
    eps = 1e-6
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        flat = p[k].reshape(-1)
        for i in rng.choice(flat.size, size=min(flat.size, 12), replace=False):
            old = flat[i]
            flat[i] = old + eps
            lp = M.loss(p, X, y, tau=tau)
            flat[i] = old - eps
            lm = M.loss(p, X, y, tau=tau)
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[k].reshape(-1)[i]) <= 1e-6 + 1e-5 * abs(fd), (k, i)


def test_bf16_mode_is_close_to_float64():
    """The rounding points only perturb: a dropped term or a sign error would not be."""
    rng = np.random.default_rng(3)
    p = M.init(64, 256, 3)
    X = rng.standard_normal((64, 64))
    y = rng.standard_normal(64)
    z64, c64 = M.forward(p, X, bf=False)
    z16, c16 = M.forward(p, X, bf=True)
    assert np.max(np.abs(z16 - z64)) < 0.02 * np.max(np.abs(z64))
    g = M.loss_grad(z64, y, bf=False)[1]  # one upstream gradient for both
    g64 = M.backward(c64, g, bf=False)
    g16 = M.backward(c16, g, bf=True)
    for k in g64:
        assert np.linalg.norm(g16[k] - g64[k]) < 0.1 * np.linalg.norm(g64[k]) + 1e-12, k


def test_pinball_known_values():
    """rho_0.9 of y - z: under-prediction costs 0.9 per unit, over-prediction 0.1;
    the minimiser over a constant is the 0.9-quantile of y."""
    z = np.array([0.0, 0.0, 2.0])
    y = np.array([1.0, -1.0, 2.0])
    l, g = M.loss_grad(z, y, bf=False, tau=0.9)
    assert np.isclose(l, (0.9 + 0.1 + 0.0) / 3)
    assert np.allclose(g, np.array([-0.9, 0.1, -0.9]) / 3)
    ys = np.arange(100.0)
    cs = np.linspace(-5, 105, 1101)
    best = cs[np.argmin([M.loss_grad(np.full(100, c), ys, bf=False, tau=0.9)[0] for c in cs])]
    assert 89.0 <= best <= 90.0


def test_coverage_label():
    order = np.array([5, 2, 7, 0, 1, 3, 4, 6])
    assert M.coverage_label(order, [2]) == 2
    assert M.coverage_label(order, [5, 5]) == 1
    assert M.coverage_label(order, [7, 0, 6]) == 8
