"""Chained fine-tune rounds: once a round has few active cells, the
device-resident schedule launches the following rounds together, without a
host read-back between them, each round checking on the device that its work
fits the tables and launch mask the chain was captured with.  Final labels
and round counts must equal the oracle's for every chain length, including
halted chains (a round over the chain's correlation limit, or over tables
that a small work budget sized for one group batch) and the
`finetune_times` limit.

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
    # chains start once a round is small, where the tail kernel would take over
    e.set_finetune_tail(0)
    yield e
    e.set_finetune_chain(0)
    e.close()


_CACHE = {}


def _case(name, thre):
    """Dataset, coded fine-tune inputs, oracle mask and oracle fine-tune."""
    key = (name, thre)
    if key not in _CACHE:
        from cusingler_amd import engine, synth
        from oracle import oracle as O
        ds = {
            "tiny_dirichlet": lambda: synth.tiny(n_ref_per_label=9, n_labels=4, n_genes=400,
                                                 n_qry=40, props="dirichlet", seed=5),
            "tiny_nan": lambda: synth.with_empty_rows(
                synth.tiny(n_ref_per_label=10, n_labels=3, n_genes=320, n_qry=30, seed=21),
                [3, 17], [5, 11]),
            "small": lambda: synth.generate("small"),
            "pancreas_2k": lambda: synth.generate("pancreas", n_qry=2000),
        }[name]()
        prep = engine.Prep(ds, 8)
        orc = O.Oracle(ds, 8)
        _, om, _ = orc.score(0, ds.qry_height, 0.8, thre, O.VAR_SEQ)
        _CACHE[key] = (ds, prep.generate(1), om, orc)
    return _CACHE[key]


def _oracle_ft(name, thre, times):
    from oracle import oracle as O
    key = (name, thre, "ft", times)
    if key not in _CACHE:
        ds, _, om, orc = _case(name, thre)
        _CACHE[key] = orc.finetune(0, ds.qry_height, om, 0.8, thre, times, O.VAR_SEQ)
    return _CACHE[key]


def _run(eng, name, thre, times, chain, path="plain"):
    ds, c1, om, _ = _case(name, thre)
    eng.copyin(c1, om)
    eng.set_finetune_path(path)
    eng.set_finetune_chain(chain)
    try:
        final, rounds = eng.finetune(0.8, thre, times)
        cnt = eng.counters()
    finally:
        eng.set_finetune_chain(0)
        eng.set_finetune_path("plain")
    eng.destroy()
    return final, rounds, cnt


@pytest.mark.parametrize("chain", [0, 1, 3, 16])
@pytest.mark.parametrize("name,thre", [("tiny_dirichlet", 0.05), ("tiny_nan", 0.05),
                                       ("small", 0.05), ("small", 0.3), ("pancreas_2k", 0.05)])
def test_chained_rounds_match_oracle(eng, name, thre, chain):
    of, orounds = _oracle_ft(name, thre, 0)
    _say("chained fine-tune", name, thre, "chain", chain)
    final, rounds, cnt = _run(eng, name, thre, 0, chain)
    _say(f"  rounds {int(orounds.max())}, chained {cnt['ft_chained_rounds']}, "
         f"halts {cnt['ft_chain_halts']}, counted rounds {cnt['ft_rounds']}")
    assert np.array_equal(final, of), f"{(final != of).sum()} final labels differ"
    assert np.array_equal(rounds, orounds), "round counts differ"
    # every round with active cells is counted once, chained or not
    assert cnt["ft_rounds"] == int(orounds.max())
    # chains start after the first round with at most 1,024 active cells (a
    # round's active cells are those with at least that many rounds)
    small_round = orounds.max() >= 1 and (orounds == orounds.max()).sum() <= 1024
    if chain == 0 or not small_round:
        assert cnt["ft_chained_rounds"] == 0
    else:
        assert cnt["ft_chained_rounds"] > 0


@pytest.mark.parametrize("name", ["small", "pancreas_2k"])
def test_chained_rounds_exact_form(eng, name):
    of, orounds = _oracle_ft(name, 0.05, 0)
    _say("chained exact fine-tune", name)
    final, rounds, cnt = _run(eng, name, 0.05, 0, 16, path="exact_int8")
    assert np.array_equal(final, of), f"{(final != of).sum()} final labels differ"
    assert np.array_equal(rounds, orounds), "round counts differ"
    assert cnt["ft_chained_rounds"] > 0


@pytest.mark.parametrize("times", [1, 2, 3])
def test_chained_rounds_times_limit(eng, times):
    of, orounds = _oracle_ft("small", 0.05, times)
    _say("chained fine-tune, times", times)
    final, rounds, cnt = _run(eng, "small", 0.05, times, 16)
    assert np.array_equal(final, of)
    assert np.array_equal(rounds, orounds)
    assert rounds.max() <= times


@pytest.mark.parametrize("path", ["plain", "exact_int8"])
def test_chained_rounds_halt(eng, path):
    """A chained round whose correlations exceed the chain's limit halts the
    chain with its state intact, and the host runs it unchained: with a limit
    of one correlation every chained round with work halts.  Labels and round
    counts are unchanged."""
    of, orounds = _oracle_ft("small", 0.05, 0)
    ds, c1, om, _ = _case("small", 0.05)
    _say("chained fine-tune, every chain halted", path)
    eng.copyin(c1, om)
    eng.set_finetune_path(path)
    eng.set_finetune_chain(16, 1)
    try:
        final, rounds = eng.finetune(0.8, 0.05, 0)
        cnt = eng.counters()
    finally:
        eng.set_finetune_chain(0)
        eng.set_finetune_path("plain")
    eng.destroy()
    _say(f"  rounds {int(orounds.max())}, chained {cnt['ft_chained_rounds']}, "
         f"halts {cnt['ft_chain_halts']}, counted rounds {cnt['ft_rounds']}")
    assert np.array_equal(final, of)
    assert np.array_equal(rounds, orounds)
    assert cnt["ft_rounds"] == int(orounds.max())
    assert cnt["ft_chain_halts"] > 0


def test_chained_rounds_after_batched_rounds(eng):
    """A small work budget runs the large rounds in group batches, whose
    tables hold one batch; chained rounds that need more halt.  Labels and
    round counts are unchanged."""
    of, orounds = _oracle_ft("small", 0.3, 0)
    ds, c1, om, _ = _case("small", 0.3)
    _say("chained fine-tune with a 256 KiB budget")
    eng.copyin(c1, om)
    eng.set_finetune_chain(16)
    before = eng.set_finetune_budget(1 << 18)
    try:
        final, rounds = eng.finetune(0.8, 0.3, 0)
        cnt = eng.counters()
        batches = eng.set_finetune_budget(8 << 30) - before
    finally:
        eng.set_finetune_budget(8 << 30)
        eng.set_finetune_chain(0)
    eng.destroy()
    _say(f"  batches {batches}, chained {cnt['ft_chained_rounds']}, "
         f"halts {cnt['ft_chain_halts']}")
    assert np.array_equal(final, of)
    assert np.array_equal(rounds, orounds)
    assert batches > 1
