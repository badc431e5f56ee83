"""The libstdc++ std::sort / std::minmax_element emulation the device uses for
the reference's percentile and label rules, checked against the real library
(the oracle calls std::sort and std::minmax_element) on NaN-laden inputs."""

import numpy as np

from cusingler_amd import engine
from oracle import oracle as O


def test_percentile_emulation_bit_exact_with_nan():
    rng = np.random.default_rng(7)
    n_cases = 0
    for _ in range(6000):
        n = int(rng.integers(1, 400))
        a = rng.random(n).astype(np.float32)
        if rng.random() < 0.7:
            k = int(rng.integers(1, max(2, n // 3)))
            a[rng.choice(n, size=min(k, n), replace=False)] = np.nan
        if rng.random() < 0.3:
            a = (np.round(a * 4) / 4).astype(np.float32)
        q = float(rng.choice([0.8, 0.5, 0.0, 1.0, 0.33]))
        try:
            h = engine.host_percentile(a, q)
        except engine.CsgError:
            continue  # the library would read out of bounds: no result claimed
        r = O.percentile(a, q)
        assert np.float32(h).tobytes() == np.float32(r).tobytes(), (n, q)
        n_cases += 1
    assert n_cases > 5000


def test_minmax_emulation():
    rng = np.random.default_rng(11)
    for _ in range(3000):
        n = int(rng.integers(1, 60))
        a = rng.integers(0, 5, size=n).astype(np.float32)
        if rng.random() < 0.5:
            a[rng.integers(0, n)] = np.nan
        assert engine.host_minmax(a) == O.minmax(a)
