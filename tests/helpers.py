"""Shared helpers of the tests (datasets, comparisons)."""

from __future__ import annotations

import numpy as np


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bits_equal(a, b, what=""):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ba, bb = bits(a), bits(b)
    # NaN payloads may differ; compare NaN-ness separately
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), f"{what}: NaN positions differ ({na.sum()} vs {nb.sum()})"
    diff = (ba != bb) & ~na
    if diff.any():
        idx = np.argwhere(diff)[:5]
        raise AssertionError(f"{what}: {int(diff.sum())} values differ in bits, e.g. at "
                             f"{idx.tolist()}: {a[diff][:5]} vs {b[diff][:5]}")
