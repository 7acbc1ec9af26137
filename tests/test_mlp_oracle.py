"""CPU tests pinning the prescreen oracle (oracle/mlp_oracle.py): finite
differences of its float64 loss, bf16 rounding against float32 bit patterns,
the pair index and the RBF feature definition."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    r = M.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie -> even
    assert r[2] == 1.0 + 2 ** -6           # tie -> even (up)
    assert r[3] == -2.5
    bits = M.bf16_bits(np.float32(1.0))
    assert int(bits) == 0x3F80
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.all(np.abs(M.bf16_round(x) - x) <= np.abs(x) * 2 ** -8)
    assert np.any(np.abs(M.bf16_round(x) - x) > np.abs(x) * 2 ** -9)


def test_pair_index_bijection():
    idx = {int(M.pair_index(a, b)) for a in range(20) for b in range(a, 20)}
    assert idx == set(range(210))
    assert M.pair_index(3, 5) == M.pair_index(5, 3)


def test_features_closed_form():
    # one pair at distance exactly mu_3 contributes 1 to bin 3 and
    # exp(-(0.8 m)^2 / (2 0.8^2)) to bin 3 + m
    d = M.RBF_START + 3 * M.RBF_STEP
    f = M.features([[0, 0, 0]], [4], [[d, 0, 0]], [9]).reshape(210, 16)
    row = f[M.pair_index(4, 9)]
    assert row[3] == pytest.approx(1.0)
    for m in (1, 2):
        assert row[3 + m] == pytest.approx(np.exp(-m * m / 2.0))
    assert f.sum() == pytest.approx(row.sum())
    f2 = M.features([[0, 0, 0]], [4], [[14.5, 0, 0]], [9])
    assert f2.sum() == 0.0                 # beyond the 14 A feature cutoff


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    d, h, B = 12, 5, 7
    n = d * h + h + h * h + h + h * h + h + h + 1
    p = rng.normal(scale=0.4, size=n)
    x = rng.normal(size=(B, d))
    t = rng.normal(size=B)
    loss, g, _ = M.loss_grad(p, x, t, d, h, bf16=False)
    eps = 1e-6
    for i in rng.choice(n, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_grad(pp, x, t, d, h, bf16=False)[0] -
              M.loss_grad(pm, x, t, d, h, bf16=False)[0]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8)
