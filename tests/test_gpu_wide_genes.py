"""Large gene sets on the GPU: score gene sets of about 2,900 and 3,750 genes
and fine-tune universes of about 6,900 and 10,800 genes, whose first rounds
rank and correlate gene subsets of 2,900-3,750 genes (the reference ranks and
correlates any subset: src/cusingler.cu:1158-1249).  The FP32 correlation
tiles then walk their genes in LDS stages.  Scores, labels, final labels and
round counts against the CPU oracle, on both fine-tune schedules, in FP32 and
exact form, with the fused form on and off.

Every case prints its name before touching the GPU."""

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


def _data(name):
    from cusingler_amd import synth
    # label_sigma: every gene's mean differs between labels, so label pairs
    # differ in thousands of genes; 16-bit value codes in "wide4k"
    cfg = {
        "wide3k": synth.SynthConfig("wide3k", 32 * 6, 60, 16000, 0, 0, 32, 20, 2.0, 6.0, 0.3,
                                    seed=3, label_sigma=1.0, marker_pool=1.0),
        "wide4k": synth.SynthConfig("wide4k", 40 * 5, 60, 20000, 0, 0, 40, 20, 2.0, 6.0, 0.4,
                                    seed=3, label_sigma=2.0, marker_pool=1.0),
    }[name]
    return synth.generate(cfg)


_CACHE = {}


def _case(name):
    """Dataset, coded steps and the oracle's results: the score step, and the
    fine-tune from masks that hold every label (first round over the whole
    pair union) or random label subsets, with thresholds 0.05 and 0.3."""
    if name not in _CACHE:
        from cusingler_amd import engine
        from oracle import oracle as O
        ds = _data(name)
        prep = engine.Prep(ds, 8)
        c0, c1 = prep.generate(0), prep.generate(1)
        orc = O.Oracle(ds, 8)
        sc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
        L = c1.ct_num
        rng = np.random.default_rng(7)
        mask = np.ones((ds.qry_height, L), np.float32)
        for q in range(1, ds.qry_height, 2):
            mask[q] = 0
            mask[q, rng.choice(L, size=int(rng.integers(2, L + 1)), replace=False)] = 1
        ft = {thre: orc.finetune(0, ds.qry_height, mask, 0.8, thre, 0, O.VAR_SEQ)
              for thre in (0.05, 0.3)}
        _CACHE[name] = (ds, c0, c1, sc, mask, ft)
    return _CACHE[name]


@pytest.mark.parametrize("name", ["wide3k", "wide4k"])
def test_wide_score_step(eng, name):
    ds, c0, c1, (of, om, osc), _, _ = _case(name)
    assert c0.width > 2800 and c1.width > 6000
    for path in ("plain", "exact_int8"):
        _say("wide score step", name, path, "G_s", c0.width)
        eng.copyin_score(c0)
        first, mask, scores = eng.get_label(0.8, 0.05, path)
        eng.destroy_score()
        assert np.array_equal(first, of), (path, int((first != of).sum()))
        assert np.array_equal(mask, om), path
        if path == "plain":
            assert_bits_equal(scores, osc, f"{name} plain scores")


@pytest.mark.parametrize("schedule", ["device", "host"])
@pytest.mark.parametrize("name", ["wide3k", "wide4k"])
def test_wide_finetune(eng, name, schedule):
    ds, c0, c1, _, mask, ft = _case(name)
    codes = max(int(c1.ref_data.max()), int(c1.qry_data.max()))
    try:
        for path in ("plain", "exact_int8"):
            for fuse in ((0, 4) if schedule == "device" else (4,)):
                for thre in (0.05, 0.3):
                    _say("wide fine-tune", name, schedule, path, "fuse", fuse, "thre", thre,
                         "G_f", c1.width, "codes", codes)
                    eng.copyin(c1, mask)
                    eng.set_finetune_path(path)
                    eng.set_finetune_schedule(schedule)
                    eng.set_finetune_fuse(fuse)
                    final, rounds = eng.finetune(0.8, thre, 0)
                    eng.destroy()
                    of, orounds = ft[thre]
                    assert np.array_equal(final, of), (path, fuse, thre, int((final != of).sum()))
                    assert np.array_equal(rounds, orounds), (path, fuse, thre)
    finally:
        eng.set_finetune_fuse(4)
        eng.set_finetune_schedule("device")
        eng.set_finetune_path("plain")


@pytest.mark.parametrize("gs", [2000, 3000, 4096, 5000, 9000])
def test_finetune_chunk_tree_staged(eng, gs):
    """The FP32 chunk-tree tiles over gene subsets beyond one LDS stage, and
    with 64-gene stages, bit-equal to the host's 128-chunk + tree order."""
    _say("ft chunk tree staged", gs)
    assert eng.selftest_ft(gs) == 0


def test_wide_proposal(eng):
    """The learned proposal on a score gene set above 1,920 genes (the
    grouped contraction's three-partial 64 x 128 blocks): every scored
    quantile equals the full exact path's, or one of the two re-scored the
    pair in the reference's FP32 order, and the labels agree with the
    oracle's."""
    import dataclasses
    from cusingler_amd import engine, synth
    from oracle import oracle as O
    cfg = dataclasses.replace(
        synth.SynthConfig("wide3k", 32 * 6, 600, 16000, 0, 0, 32, 20, 2.0, 6.0, 0.3, seed=3,
                          label_sigma=1.0, marker_pool=1.0))
    ds = synth.generate(cfg)
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    orc = O.Oracle(ds, 8)
    of, om, osc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    ofin, _ = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    assert c0.width > 1920
    _say("wide proposal: train, G_s", c0.width)
    eng.copyin_score(c0)
    eng.mlp_init(64, seed=5)
    eng.mlp_train(ofin.astype(np.int32), epochs=20, batch=128, lr=2e-3, seed=1)
    _say("wide proposal: full exact path")
    fe, me, se = eng.get_label(0.8, 0.05, "exact_int8")
    _say("wide proposal: proposal path")
    eng.set_proposal(4, 0.2)
    try:
        fp, mp, sp = eng.get_label(0.8, 0.05, "exact_int8")
        cnt = eng.counters()
    finally:
        eng.set_proposal(0)
    eng.copyin(c1)
    finp, _ = eng.finetune(0.8, 0.05, 0)
    eng.destroy()
    eng.destroy_score()
    _say("wide proposal cells", cnt["prop_cells"], "scored pairs", cnt["prop_scored_pairs"],
         "of", ds.qry_height * c0.ct_num)
    assert cnt["prop_cells"] > 0
    assert cnt["prop_scored_pairs"] < ds.qry_height * c0.ct_num
    fin = np.isfinite(sp)
    u, ue, uo = sp.view(np.uint32), se.view(np.uint32), osc.view(np.uint32)
    odd = fin & (u != ue)
    # where the two differ, one of them re-scored the pair in FP32 order
    assert np.all((ue[odd] == uo[odd]) | (u[odd] == uo[odd])), int(odd.sum())
    assert np.array_equal(fe, of) and np.array_equal(me, om)
    assert float((fp == of).mean()) >= 0.99 and float((finp == ofin).mean()) >= 0.99
