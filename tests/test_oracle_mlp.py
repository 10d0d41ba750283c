"""Pins of the oracle's network (CPU): the analytic gradient against central
finite differences, and the structure of the forward pass on closed forms."""
from __future__ import annotations

import numpy as np

import oracle

NP = 34 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3
OFF_W1, OFF_B1 = 0, 34 * 64
OFF_W2 = OFF_B1 + 64
OFF_B2 = OFF_W2 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 3 * 64


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(NP) * scale).astype(np.float32)
    p[OFF_B1:OFF_B1 + 64] = np.abs(p[OFF_B1:OFF_B1 + 64])  # keep units alive
    return p


def _data(n=40, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 34)).astype(np.float32)
    a = rng.random((n, 3)).astype(np.float32)
    t = rng.random((n, 3)).astype(np.float32)
    return x, a, t


def test_gradient_matches_central_differences():
    p = _params()
    x, a, t = _data()
    loss, g = oracle.mlp_loss_grad(p, x, a, t)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(np.arange(OFF_W1, OFF_B1), 20), np.arange(OFF_B1, OFF_B1 + 5),
                          rng.choice(np.arange(OFF_W2, OFF_B2), 20), np.arange(OFF_B2, OFF_B2 + 5),
                          rng.choice(np.arange(OFF_W3, OFF_B3), 20), np.arange(OFF_B3, NP)])
    h = 1e-3
    for i in idx:
        pp = p.astype(np.float64).copy()
        pp[i] += h
        lp, _ = oracle.mlp_loss_grad(pp.astype(np.float32), x, a, t)
        pp[i] -= 2 * h
        lm, _ = oracle.mlp_loss_grad(pp.astype(np.float32), x, a, t)
        # float32 parameters: the step actually taken is the rounded one
        hp = float(np.float32(p[i] + h)) - float(p[i])
        hm = float(p[i]) - float(np.float32(p[i] - h))
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - g[i]) <= 2e-4 * max(1.0, abs(g[i])) + 2e-6, (i, fd, g[i])


def test_forward_closed_form_linear_regime():
    # all pre-activations positive -> the network is the affine map W3 W2 W1 x + ...
    rng = np.random.default_rng(3)
    p = np.zeros(NP, np.float32)
    W1 = rng.random((64, 34)).astype(np.float32) * 0.1
    W2 = rng.random((64, 64)).astype(np.float32) * 0.1
    W3 = rng.standard_normal((3, 64)).astype(np.float32) * 0.1
    b1 = rng.random(64).astype(np.float32)
    b2 = rng.random(64).astype(np.float32)
    b3 = rng.standard_normal(3).astype(np.float32)
    p[OFF_W1:OFF_B1] = W1.reshape(-1)
    p[OFF_B1:OFF_W2] = b1
    p[OFF_W2:OFF_B2] = W2.reshape(-1)
    p[OFF_B2:OFF_W3] = b2
    p[OFF_W3:OFF_B3] = W3.reshape(-1)
    p[OFF_B3:] = b3
    x = rng.random((10, 34)).astype(np.float32)
    W1d, W2d, W3d = W1.astype(np.float64), W2.astype(np.float64), W3.astype(np.float64)
    want = (W3d @ (W2d @ (W1d @ x.T.astype(np.float64) + b1[:, None]) + b2[:, None]) + b3[:, None]).T
    assert np.allclose(oracle.mlp_forward(p, x), want, rtol=1e-12, atol=1e-12)


def test_relu_cuts_negative_units_and_bf16_mode_is_close():
    p = _params(4)
    x, a, t = _data(64, 5)
    y = oracle.mlp_forward(p, x)
    yq = oracle.mlp_forward(p, x, bf16=True)
    assert np.abs(y - yq).max() < 5e-2 * max(1.0, np.abs(y).max())
    l, g = oracle.mlp_loss_grad(p, x, a, t)
    lq, gq = oracle.mlp_loss_grad(p, x, a, t, bf16=True)
    assert abs(l - lq) < 2e-2 * l
    assert np.linalg.norm(g - gq) < 5e-2 * np.linalg.norm(g)
    # zero albedo -> zero loss and zero gradient
    l0, g0 = oracle.mlp_loss_grad(p, x, np.zeros_like(a), t)
    assert l0 == 0.0 and not g0.any()
