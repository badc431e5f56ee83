"""Pins the proposal network's CPU oracle (oracle/mlp_oracle.py).

Its FP64 gradient is checked against central finite differences of its own
loss at parameters of every block (so a dropped term, a wrong sign or a
transposed operand in the backward pass fails), its BF16 rounding against
exhaustive properties of round-to-nearest-even, and the loss against the
closed form log(L) of an all-zero network.
"""

import numpy as np
import pytest

from oracle import mlp_oracle as M


def _setup(seed=0, n=7, G=13, H=8, L=5):
    rng = np.random.default_rng(seed)
    d = M.Dims(G=G, Kp=16, H=H, L=L, Lp=8)
    w = rng.normal(0, 0.5, d.n_params)
    # inert padding: input rows >= G and output columns >= L are zero
    parts = M.unpack(w, d)
    parts[0][G:, :] = 0
    parts[4][:, L:] = 0
    parts[5][L:] = 0
    w = M.pack(parts)
    x = rng.uniform(-1, 1, (n, G))
    y = rng.integers(0, L, n)
    return d, w, x, y


def test_gradient_matches_finite_differences():
    d, w, x, y = _setup()
    loss, g = M.loss_and_grad(w, x, y, d, bf16=False)
    rng = np.random.default_rng(1)
    offs = np.cumsum([0] + d.sizes)
    idx = []
    for b in range(6):
        lo, hi = offs[b], offs[b + 1]
        # only live parameters (padding gets zero gradient by construction)
        cand = [i for i in range(lo, hi) if g[i] != 0.0]
        idx += list(rng.choice(cand, size=min(6, len(cand)), replace=False))
    fd = M.finite_difference(w, x, y, d, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


def test_padding_parameters_get_zero_gradient():
    d, w, x, y = _setup(seed=3)
    _, g = M.loss_and_grad(w, x, y, d, bf16=False)
    gW1, _, _, _, gW3, gb3 = M.unpack(g, d)
    assert np.all(gW1[d.G:, :] == 0)
    assert np.all(gW3[:, d.L:] == 0)
    assert np.all(gb3[d.L:] == 0)


def test_zero_network_loss_is_log_L():
    d, w, x, y = _setup(seed=4)
    loss, g = M.loss_and_grad(np.zeros_like(w), x, y, d, bf16=False)
    assert loss == pytest.approx(np.log(d.L), rel=1e-12)
    # with zero weights only the output bias receives a gradient: mean(p - onehot)
    gb3 = M.unpack(g, d)[5][:d.L]
    expect = np.full(d.L, 1.0 / d.L) - np.bincount(y, minlength=d.L) / len(y)
    np.testing.assert_allclose(gb3, expect, atol=1e-15)


def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(5)
    v = rng.normal(0, 10, 20000).astype(np.float32)
    r = M.bf16_round(v)
    # representable: low 16 bits clear
    assert np.all((r.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)
    # nearest: no other bf16 neighbour is closer
    u = r.astype(np.float32).view(np.uint32)
    up = (u + 0x10000).view(np.float32).astype(np.float64)
    dn = (u - 0x10000).view(np.float32).astype(np.float64)
    err = np.abs(r - v)
    assert np.all(err <= np.abs(up - v)) and np.all(err <= np.abs(dn - v))
    # ties go to the even mantissa
    ties = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8], dtype=np.float32)
    np.testing.assert_array_equal(M.bf16_round(ties), [1.0, 1.0 + 4 * 2 ** -8])


def test_bf16_mode_stays_close_to_fp64():
    d, w, x, y = _setup(seed=6)
    l64, g64 = M.loss_and_grad(w, x, y, d, bf16=False)
    l16, g16 = M.loss_and_grad(w, x, y, d, bf16=True)
    assert abs(l16 - l64) < 2e-2 * abs(l64)
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)


def test_planted_transpose_is_caught():
    """A backward pass that multiplies by W2 instead of W2^T (a plausible
    mistake) disagrees with finite differences."""
    d, w, x, y = _setup(seed=7, H=8)
    W1, b1, W2, b2, W3, b3 = M.unpack(w, d)
    # make W2 clearly non-symmetric
    W2 += np.triu(np.ones_like(W2))
    w = M.pack([W1, b1, W2, b2, W3, b3])

    def wrong(w):
        W1, b1, W2, b2, W3, b3 = M.unpack(w, d)
        xp = np.zeros((x.shape[0], d.Kp))
        xp[:, :d.G] = x
        z1 = xp @ W1 + b1
        h1 = M.gelu(z1)
        z2 = h1 @ W2 + b2
        h2 = M.gelu(z2)
        lg = (h2 @ W3 + b3)[:, :d.L]
        p = np.exp(lg - lg.max(1, keepdims=True))
        p /= p.sum(1, keepdims=True)
        dl = np.zeros((x.shape[0], d.Lp))
        dl[:, :d.L] = p
        dl[np.arange(len(y)), y] -= 1
        dl /= len(y)
        dz2 = (dl @ W3.T) * M.gelu_grad(z2)
        dz1 = (dz2 @ W2) * M.gelu_grad(z1)  # planted: W2 instead of W2.T
        return xp.T @ dz1

    gW1_wrong = wrong(w)
    _, g = M.loss_and_grad(w, x, y, d, bf16=False)
    gW1 = M.unpack(g, d)[0]
    idx = [i * d.H + o for i in range(3) for o in range(3)]
    fd = M.finite_difference(w, x, y, d, idx)
    np.testing.assert_allclose(gW1.ravel()[idx], fd, rtol=1e-6, atol=1e-9)
    assert not np.allclose(gW1_wrong.ravel()[idx], fd, rtol=1e-3, atol=1e-6)
