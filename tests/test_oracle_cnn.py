"""Pins of the triage-CNN oracle: its gradient against central finite
differences of its own loss (FP64, small input), a closed form (zero
weights), and the BF16 rounding rule (round to nearest even)."""
import numpy as np
import pytest

import oracle


def _init(H, W, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = np.zeros(oracle.cnn_param_count(H, W))
    for name, a, b in oracle.cnn_param_slices(H, W):
        fan = {"w1": 18, "w2": 144, "w3": 288, "fw1": 64 * (H // 8) * (W // 8), "fw2": 64}.get(name)
        p[a:b] = rng.standard_normal(b - a) * (scale / np.sqrt(fan) * 3 if fan else 0.1)
    return p


def test_param_layout():
    H, W = 16, 32
    sl = oracle.cnn_param_slices(H, W)
    assert sl[-1][2] == oracle.cnn_param_count(H, W)
    assert [n for n, _, _ in sl] == ["w1", "b1", "w2", "b2", "w3", "b3", "fw1", "fb1", "fw2", "fb2"]
    assert sl[6][2] - sl[6][1] == 64 * 2 * 4 * 64


def test_gradient_matches_finite_differences():
    H, W, B = 8, 16, 3
    rng = np.random.default_rng(1)
    x = rng.standard_normal((B, 2, H, W))
    labels = np.array([0, 1, 1], np.int32)
    p = _init(H, W, 2)
    loss, g, _ = oracle.cnn_loss_grad(x, labels, p)
    assert np.isfinite(loss) and loss > 0
    idx = []
    for name, a, b in oracle.cnn_param_slices(H, W):
        idx += list(rng.choice(np.arange(a, b), size=min(6, b - a), replace=False))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (oracle.cnn_loss_grad(x, labels, pp, grads=False)[0] -
              oracle.cnn_loss_grad(x, labels, pm, grads=False)[0]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-8), i


def test_single_class_logits_closed_form():
    # all weights zero: logits are the fc2 bias, loss = softplus of the margin
    H, W = 8, 8
    p = np.zeros(oracle.cnn_param_count(H, W))
    p[-2:] = [0.5, -0.25]
    x = np.random.default_rng(3).standard_normal((2, 2, H, W))
    loss, g, logits = oracle.cnn_loss_grad(x, np.array([0, 1], np.int32), p)
    np.testing.assert_allclose(logits, [[0.5, -0.25], [0.5, -0.25]])
    l0 = np.log1p(np.exp(-0.75))
    l1 = np.log1p(np.exp(0.75))
    assert loss == pytest.approx((l0 + l1) / 2, rel=1e-14)
    s = 1 / (1 + np.exp(-0.75))  # softmax probability of class 0
    # d loss / d logits: sample 0 (label 0) [s - 1, 1 - s], sample 1 (label 1) [s, -s]
    np.testing.assert_allclose(g[-2:], [(2 * s - 1) / 2, (1 - 2 * s) / 2], rtol=1e-12)
    assert not g[:-2].any()


def test_bf16_rounding_rule():
    assert oracle.bf16_round(1.0) == 1.0
    assert oracle.bf16_round(1 + 2 ** -8) == 1.0  # tie -> even
    assert oracle.bf16_round(1 + 3 * 2 ** -8) == 1 + 2 ** -6  # tie -> even (up)
    assert oracle.bf16_round(1 + 2 ** -8 + 2 ** -20) == 1 + 2 ** -7
    assert oracle.bf16_round(-3.14159) == -3.140625


def test_bf16_mode_close_to_fp64():
    H, W, B = 16, 32, 2
    rng = np.random.default_rng(4)
    x = rng.standard_normal((B, 2, H, W))
    p = _init(H, W, 5)
    l64, g64, _ = oracle.cnn_loss_grad(x, np.array([1, 0], np.int32), p)
    l16, g16, _ = oracle.cnn_loss_grad(x, np.array([1, 0], np.int32), p, bf16=True)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2
