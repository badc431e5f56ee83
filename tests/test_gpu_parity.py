"""Parity of the HIP engine with the CPU oracle (bit-exact scores and labels).

Every case prints its name before touching the GPU so that a failing run
names it."""

import sys

import numpy as np
import pytest

from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from cusingler_amd import engine
    e = engine.Engine(0)
    # every round one launch sequence at a time (the tail kernel has tests of its own)
    e.set_finetune_tail(0)
    yield e
    e.close()


def _datasets():
    from cusingler_amd import synth
    return {
        "tiny": lambda: synth.tiny(),
        "tiny_filtered": lambda: synth.tiny(n_ref_only=7, n_qry_only=5, seed=11),
        "tiny_dirichlet": lambda: synth.tiny(n_ref_per_label=9, n_labels=4, n_genes=400, n_qry=40,
                                             props="dirichlet", seed=5),
        "small": lambda: synth.generate("small"),
        "pancreas_2k": lambda: synth.generate("pancreas", n_qry=2000),
        # empty reference and query cells: constant ranks, NaN correlations
        "tiny_nan": lambda: synth.with_empty_rows(
            synth.tiny(n_ref_per_label=10, n_labels=3, n_genes=320, n_qry=30, seed=21),
            [3, 17], [5, 11]),
    }


ALL = ["tiny", "tiny_filtered", "tiny_dirichlet", "tiny_nan", "small", "pancreas_2k"]


@pytest.mark.parametrize("name", ALL)
def test_score_step_plain_bit_exact(eng, name):
    from cusingler_amd import engine
    from oracle import oracle as O
    ds = _datasets()[name]()
    prep = engine.Prep(ds, 8)
    c0 = prep.generate(0)
    _say("score plain", name, ds.qry_height, ds.ref_height, c0.width)
    eng.copyin_score(c0)
    first, mask, scores = eng.get_label(0.8, 0.05, "plain")
    eng.destroy_score()
    orc = O.Oracle(ds, 8)
    of, om, osc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    assert_bits_equal(scores, osc, f"{name} per-label scores")
    assert np.array_equal(first, of), f"{name} first labels differ"
    assert np.array_equal(mask, om), f"{name} candidate masks differ"


@pytest.mark.parametrize("schedule", ["device", "host"])
@pytest.mark.parametrize("name", ALL)
def test_finetune_plain_matches_oracle(eng, name, schedule):
    from cusingler_amd import engine
    from oracle import oracle as O
    ds = _datasets()[name]()
    prep = engine.Prep(ds, 8)
    orc = O.Oracle(ds, 8)
    _, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    of, orounds = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    c1 = prep.generate(1)
    _say("finetune plain", name, schedule, ds.qry_height, c1.width)
    eng.copyin(c1, om)
    eng.set_finetune_schedule(schedule)
    try:
        final, rounds = eng.finetune(0.8, 0.05, 0)
    finally:
        eng.set_finetune_schedule("device")
    eng.destroy()
    assert np.array_equal(final, of), f"{name}: {(final != of).sum()} final labels differ"
    assert np.array_equal(rounds, orounds), f"{name}: round counts differ"


@pytest.mark.parametrize("schedule", ["device", "host"])
def test_finetune_times_limit(eng, schedule):
    from cusingler_amd import engine, synth
    from oracle import oracle as O
    ds = synth.generate("small")
    prep = engine.Prep(ds, 8)
    orc = O.Oracle(ds, 8)
    _, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    # a wide threshold keeps many candidates so the limit matters
    _, om2, _ = orc.score(0, ds.qry_height, 0.8, 0.3, O.VAR_SEQ)
    for times in (1, 2):
        of, _ = orc.finetune(0, ds.qry_height, om2, 0.8, 0.3, times, O.VAR_SEQ)
        _say("finetune times", schedule, times)
        eng.copyin(prep.generate(1), om2)
        eng.set_finetune_schedule(schedule)
        try:
            final, rounds = eng.finetune(0.8, 0.3, times)
        finally:
            eng.set_finetune_schedule("device")
        eng.destroy()
        assert np.array_equal(final, of)
        assert rounds.max() <= times


def test_pipeline_end_to_end(eng):
    """train -> score -> finetune through the device mask, against the oracle."""
    from cusingler_amd import engine, synth
    from oracle import oracle as O
    ds = synth.generate("small", n_qry=600)
    _say("pipeline small 600")
    ann = engine.annotate(ds, "plain", engine=eng)
    orc = O.Oracle(ds, 8)
    of, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    ofin, _ = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    assert np.array_equal(ann.first, of)
    assert np.array_equal(ann.final, ofin)


def test_pybind_run_matches_oracle():
    from cusingler import cusingler
    from cusingler_amd import synth
    from oracle import oracle as O
    ds = synth.tiny(n_ref_per_label=8, n_labels=3, n_genes=320, n_qry=30, seed=3)
    _say("pybind run tiny")
    res = cusingler.run(4, 0, ds.ref_height, ds.ref_width, ds.ref_data.tolist(),
                        ds.ref_indices.tolist(), ds.ref_indptr.astype(np.int32).tolist(),
                        ds.qry_height, ds.qry_width, ds.qry_data.tolist(), ds.qry_indices.tolist(),
                        ds.qry_indptr.astype(np.int32).tolist(), ds.codes,
                        ds.celltype.tolist(), ds.cellnames, ds.ref_genes, ds.qry_genes)
    orc = O.Oracle(ds, 4)
    of, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    ofin, _ = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    assert [r[0] for r in res] == ds.cellnames
    assert [r[1] for r in res] == [ds.codes[i] for i in of]
    assert [r[2] for r in res] == [ds.codes[i] for i in ofin]
