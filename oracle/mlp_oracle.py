"""CPU oracle for the learned prescreen (numpy, float64).

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's checks import it; the
product never does). There is no reference counterpart: the reference has no
learned component (Cargo.toml:4). What pins this oracle: finite differences
of its own float64 loss (tests/test_mlp_oracle.py), closed-form checks of
the feature definition, and bf16 rounding checked against numpy's float32
bit patterns.

Model: x -> SiLU(x W1 + b1) -> SiLU(. W2 + b2) -> SiLU(. W3 + b3) -> . w4 + b4.
With bf16=True the oracle rounds at the same points as the device path:
weights, the input rows, the hidden activations A1..A3 and the backward
pre-activation gradients dZ1..dZ3 are bf16; everything else is float64.
"""
from __future__ import annotations

import numpy as np

D, H = 3360, 256
RBF, PAIRS = 16, 210
FEAT_CUTOFF, RBF_START, RBF_STEP, RBF_SIGMA = 14.0, 2.0, 0.8, 0.8


def bf16_round(a) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    r = np.where(nan, (u >> 16).astype(np.uint32) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(a) -> np.ndarray:
    f = np.ascontiguousarray(bf16_round(a).astype(np.float32))
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def bf16_to_f64(bits) -> np.ndarray:
    return (np.asarray(bits, np.uint32) << 16).view(np.float32).astype(np.float64)


def pair_index(a, b):
    a, b = np.minimum(a, b), np.maximum(a, b)
    return a * (41 - a) // 2 + (b - a)


def features(cand_ca, cand_aa, ag_ca, ag_aa) -> np.ndarray:
    """RBF distance-histogram of one candidate (float64, before log1p)."""
    h = np.zeros((PAIRS, RBF))
    mu = RBF_START + RBF_STEP * np.arange(RBF)
    c = np.asarray(cand_ca, np.float64)
    a = np.asarray(ag_ca, np.float64)
    for i in range(len(c)):
        d = np.sqrt(((a - c[i]) ** 2).sum(1))
        for j in np.nonzero(d < FEAT_CUTOFF)[0]:
            h[pair_index(int(cand_aa[i]), int(ag_aa[j]))] += np.exp(
                -(d[j] - mu) ** 2 / (2 * RBF_SIGMA ** 2))
    return h.reshape(-1)


def unpack(p, d=D, h=H):
    o = 0
    out = {}
    for name, shape in (("W1", (d, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)),
                        ("W3", (h, h)), ("b3", (h,)), ("w4", (h,)), ("b4", (1,))):
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    assert o == len(p)
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def loss_grad(p, x, t, d=D, h=H, bf16=True):
    """Mean squared error and its gradient w.r.t. the flat parameters."""
    P = unpack(np.asarray(p, np.float64), d, h)
    r = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    Wq = {k: (r(v.astype(np.float32)) if bf16 and k in ("W1", "W2", "W3", "w4") else v)
          for k, v in P.items()}
    X = np.asarray(x, np.float64)
    B = X.shape[0]
    Z1 = X @ Wq["W1"] + P["b1"]
    A1 = r(silu(Z1).astype(np.float32)) if bf16 else silu(Z1)
    Z2 = A1 @ Wq["W2"] + P["b2"]
    A2 = r(silu(Z2).astype(np.float32)) if bf16 else silu(Z2)
    Z3 = A2 @ Wq["W3"] + P["b3"]
    A3 = r(silu(Z3).astype(np.float32)) if bf16 else silu(Z3)
    y = A3 @ Wq["w4"] + P["b4"][0]
    res = y - np.asarray(t, np.float64)
    loss = float((res ** 2).mean())
    dy = 2.0 * res / B
    g = {"w4": A3.T @ dy, "b4": np.array([dy.sum()])}
    q = (lambda v: r(v.astype(np.float32))) if bf16 else (lambda v: v)
    dZ3 = q(np.outer(dy, Wq["w4"]) * dsilu(Z3))
    g["W3"] = A2.T @ dZ3
    g["b3"] = dZ3.sum(0)
    dZ2 = q((dZ3 @ Wq["W3"].T) * dsilu(Z2))
    g["W2"] = A1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = q((dZ2 @ Wq["W2"].T) * dsilu(Z1))
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    flat = np.concatenate([g[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4",
                                                       "b4")])
    return loss, flat, y


def predict(p, x, d=D, h=H, bf16=True):
    return loss_grad(p, x, np.zeros(len(x)), d, h, bf16)[2]
