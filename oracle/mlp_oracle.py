"""CPU oracle of the learned label-proposal network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device's
training kernels; the product never imports it.

The network has no counterpart in the reference.  It maps the score step's
rank features x (doubled centred ranks / G over the score gene set) through
    z1 = x W1 + b1, h1 = gelu(z1), z2 = h1 W2 + b2, h2 = gelu(z2),
    logits = h2 W3 + b3
to softmax cross-entropy, averaged over the rows.  gelu(z) = z * Phi(z) with
the exact normal CDF.  Parameters use the engine's flat layout
[W1 (Kp x H) | b1 (H) | W2 (H x H) | b2 (H) | W3 (H x Lp) | b3 (Lp)],
row-major; only the first L logits enter the softmax.

With ``bf16=True`` the oracle rounds (round-to-nearest-even) at the points
where the device keeps BF16 values: the features, the weight matrices, h1,
h2, the output gradient and the two pre-activation gradients that feed the
next contraction; every sum is still exact to FP64.  With ``bf16=False`` it is
the plain FP64 network, whose gradient ``finite_difference`` pins.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.special import erf


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


@dataclass
class Dims:
    G: int
    Kp: int
    H: int
    L: int
    Lp: int

    @property
    def sizes(self):
        return [self.Kp * self.H, self.H, self.H * self.H, self.H, self.H * self.Lp, self.Lp]

    @property
    def n_params(self):
        return sum(self.sizes)


def unpack(w: np.ndarray, d: Dims):
    out, o = [], 0
    shapes = [(d.Kp, d.H), (d.H,), (d.H, d.H), (d.H,), (d.H, d.Lp), (d.Lp,)]
    for n, sh in zip(d.sizes, shapes):
        out.append(np.asarray(w[o:o + n], dtype=np.float64).reshape(sh))
        o += n
    return out


def pack(parts) -> np.ndarray:
    return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])


def loss_and_grad(w: np.ndarray, x: np.ndarray, y: np.ndarray, d: Dims, bf16: bool = True):
    """Mean cross-entropy over the rows of x ([n x G]) and its gradient in the
    flat parameter layout."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(w, d)
    n = x.shape[0]
    xp = np.zeros((n, d.Kp))
    xp[:, :d.G] = x
    xb = r(xp)
    W1b, W2b, W3b = r(W1), r(W2), r(W3)
    z1 = xb @ W1b + b1
    h1 = r(gelu(z1))
    z2 = h1 @ W2b + b2
    h2 = r(gelu(z2))
    logits = h2 @ W3b + b3
    lg = logits[:, :d.L]
    mx = lg.max(axis=1, keepdims=True)
    e = np.exp(lg - mx)
    s = e.sum(axis=1, keepdims=True)
    loss = float(np.mean(np.log(s[:, 0]) + mx[:, 0] - lg[np.arange(n), y]))
    dl = np.zeros((n, d.Lp))
    dl[:, :d.L] = e / s
    dl[np.arange(n), y] -= 1.0
    dl /= n
    dlb = r(dl)
    gW3 = h2.T @ dlb
    gb3 = dl.sum(axis=0)
    dz2 = (dlb @ W3b.T) * gelu_grad(z2)
    dz2b = r(dz2)
    gW2 = h1.T @ dz2b
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2b @ W2b.T) * gelu_grad(z1)
    dz1b = r(dz1)
    gW1 = xb.T @ dz1b
    gb1 = dz1.sum(axis=0)
    return loss, pack([gW1, gb1, gW2, gb2, gW3, gb3])


def loss_only(w, x, y, d: Dims) -> float:
    return loss_and_grad(w, x, y, d, bf16=False)[0]


def finite_difference(w, x, y, d: Dims, idx, h: float = 1e-5) -> np.ndarray:
    """Central differences of the FP64 loss at the parameters `idx`."""
    out = np.zeros(len(idx))
    for t, i in enumerate(idx):
        wp = np.array(w, dtype=np.float64)
        wm = wp.copy()
        wp[i] += h
        wm[i] -= h
        out[t] = (loss_only(wp, x, y, d) - loss_only(wm, x, y, d)) / (2 * h)
    return out
