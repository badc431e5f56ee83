"""The fine-tune tail kernel: once a round has few active cells, that round and
all later ones run in one cooperative launch, each cell a group of its own.
Final labels and round counts must equal the oracle's (and the per-round
schedule's) whichever round the tail starts at -- from the first round on
(every round in the tail), or after the per-round launches of the larger
rounds -- in FP32 and exact form, with NaN-carrying segments, under the
`finetune_times` limit, with 16-bit value codes and fine-tune gene sets of
thousands of genes, and for any number of blocks per CU.

Every case prints its name before touching the GPU."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from cusingler_amd import engine
    e = engine.Engine(0)
    yield e
    e.set_finetune_tail(64)
    e.close()


_CACHE = {}


def _jittered(ds, seed):
    """Every stored value scaled by its own factor in (1, 1.25): rows hold
    hundreds of distinct values, so the value codes need 16 bits."""
    import dataclasses
    rng = np.random.default_rng(seed)

    def j(d):
        return (d * (1.0 + rng.integers(1, 1000, size=d.shape[0]) / 4096.0)).astype(np.float32)

    return dataclasses.replace(ds, ref_data=j(ds.ref_data), qry_data=j(ds.qry_data))


def _dataset(name):
    from cusingler_amd import synth
    return {
        "tiny": lambda: synth.tiny(),
        "tiny_filtered": lambda: synth.tiny(n_ref_only=7, n_qry_only=5, seed=11),
        "tiny_dirichlet": lambda: synth.tiny(n_ref_per_label=9, n_labels=4, n_genes=400,
                                             n_qry=40, props="dirichlet", seed=5),
        # empty reference and query cells: NaN correlations
        "tiny_nan": lambda: synth.with_empty_rows(
            synth.tiny(n_ref_per_label=10, n_labels=3, n_genes=320, n_qry=30, seed=21),
            [3, 17], [5, 11]),
        "small": lambda: synth.generate("small"),
        "pancreas_2k": lambda: synth.generate("pancreas", n_qry=2000),
        # more than 256 value codes (16-bit codes)
        "codes16": lambda: _jittered(synth.tiny(n_ref_per_label=40, n_labels=8, n_genes=1500,
                                                n_qry=300, markers=8, nonzero=0.5, seed=33), 5),
        # fine-tune gene set of several thousand genes
        "wide3k": lambda: synth.generate(synth.SynthConfig(
            "wide3k", 32 * 6, 60, 16000, 0, 0, 32, 20, 2.0, 6.0, 0.3, seed=3, label_sigma=1.0,
            marker_pool=1.0)),
    }[name]()


def _case(name, thre):
    """Coded fine-tune inputs, the oracle's mask and fine-tune (cached)."""
    key = (name, thre)
    if key not in _CACHE:
        from cusingler_amd import engine
        from oracle import oracle as O
        ds = _dataset(name)
        prep = engine.Prep(ds, 8)
        orc = O.Oracle(ds, 8)
        _, om, _ = orc.score(0, ds.qry_height, 0.8, thre, O.VAR_SEQ)
        if name == "wide3k":
            # labels this distinct leave one candidate per cell: every label
            # for the even cells (the first round over the whole pair
            # union), random subsets for the odd ones
            L = om.shape[1]
            rng = np.random.default_rng(7)
            om = np.ones_like(om)
            for q in range(1, ds.qry_height, 2):
                om[q] = 0
                om[q, rng.choice(L, size=int(rng.integers(2, L + 1)), replace=False)] = 1
        _CACHE[key] = (ds, prep.generate(1), om, orc, {})
    return _CACHE[key]


def _oracle(name, thre, times=0):
    from oracle import oracle as O
    ds, _, om, orc, ft = _case(name, thre)
    if times not in ft:
        ft[times] = orc.finetune(0, ds.qry_height, om, 0.8, thre, times, O.VAR_SEQ)
    return ft[times]


def _run(eng, name, thre, tail, times=0, path="plain", blocks=4):
    _, c1, om, _, _ = _case(name, thre)
    eng.copyin(c1, om)
    eng.set_finetune_path(path)
    eng.set_finetune_tail(tail, blocks)
    try:
        final, rounds = eng.finetune(0.8, thre, times)
        cnt = eng.counters()
    finally:
        eng.set_finetune_tail(64)
        eng.set_finetune_path("plain")
    eng.destroy()
    return final, rounds, cnt


ALL = [("tiny", 0.05), ("tiny_filtered", 0.05), ("tiny_dirichlet", 0.05), ("tiny_nan", 0.05),
       ("small", 0.05), ("small", 0.3), ("pancreas_2k", 0.05)]


@pytest.mark.parametrize("tail", [1 << 30, 256])
@pytest.mark.parametrize("name,thre", ALL)
def test_tail_matches_oracle(eng, name, thre, tail):
    of, orounds = _oracle(name, thre)
    _say("tail fine-tune", name, thre, "tail from", tail, "active cells")
    final, rounds, cnt = _run(eng, name, thre, tail)
    _say(f"  rounds {int(orounds.max())}, tail rounds {cnt['ft_tail_rounds']}, "
         f"counted {cnt['ft_rounds']}")
    assert np.array_equal(final, of), f"{(final != of).sum()} final labels differ"
    assert np.array_equal(rounds, orounds), "round counts differ"
    assert cnt["ft_rounds"] == int(orounds.max())
    if tail > 1 << 20 and orounds.max() > 0:
        # every round in the tail
        assert cnt["ft_tail_rounds"] == int(orounds.max())


@pytest.mark.parametrize("name", ["small", "pancreas_2k", "tiny_nan"])
def test_tail_exact_form(eng, name):
    of, orounds = _oracle(name, 0.05)
    _say("tail fine-tune, exact form", name)
    final, rounds, cnt = _run(eng, name, 0.05, 1 << 30, path="exact_int8")
    assert np.array_equal(final, of), f"{(final != of).sum()} final labels differ"
    assert np.array_equal(rounds, orounds)


@pytest.mark.parametrize("times", [1, 2, 3])
def test_tail_times_limit(eng, times):
    of, orounds = _oracle("small", 0.3, times)
    _say("tail fine-tune, times", times)
    final, rounds, cnt = _run(eng, "small", 0.3, 1 << 30, times=times)
    assert np.array_equal(final, of)
    assert np.array_equal(rounds, orounds)
    assert rounds.max() <= times


@pytest.mark.parametrize("blocks", [1, 16])
def test_tail_grid_sizes(eng, blocks):
    """One block per CU (cells outnumber blocks in the first rounds) and as
    many as can be co-resident."""
    of, orounds = _oracle("small", 0.3)
    _say("tail fine-tune, blocks per CU", blocks)
    final, rounds, cnt = _run(eng, "small", 0.3, 1 << 30, blocks=blocks)
    assert np.array_equal(final, of)
    assert np.array_equal(rounds, orounds)
    assert cnt["ft_tail_rounds"] == int(orounds.max())


@pytest.mark.parametrize("name", ["codes16", "wide3k"])
def test_tail_kernel_variants(eng, name):
    """16-bit value codes and a fine-tune gene set of thousands of genes,
    against the per-round schedule and the oracle."""
    of, orounds = _oracle(name, 0.3)
    _, c1, _, _, _ = _case(name, 0.3)
    codes = max(int(c1.ref_data.max()), int(c1.qry_data.max()))
    _say("tail fine-tune", name, "G_f", c1.width, "codes", codes)
    if name == "codes16":
        assert codes > 255
    else:
        assert c1.width > 2000
    ref = _run(eng, name, 0.3, 0)
    for path in ("plain", "exact_int8"):
        final, rounds, cnt = _run(eng, name, 0.3, 1 << 30, path=path)
        assert np.array_equal(final, of), (path, int((final != of).sum()))
        assert np.array_equal(rounds, orounds), path
        assert np.array_equal(final, ref[0]) and np.array_equal(rounds, ref[1]), path
        assert cnt["ft_tail_rounds"] > 0
