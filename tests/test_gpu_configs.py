"""The two largest configurations at their stated shapes, checked against the
CPU oracle on a query subset:

- streaming: 50,000 reference cells x 30,000 genes, 100 labels as 10 parents
  x 10 children sharing 70% of markers; 2,000 query cells.  Plain scores bit
  for bit, first labels, masks, final labels and round counts on the plain and
  exact paths; the chosen form (exact INT8 + learned proposal, network trained
  on plain-method labels of 20,000 training cells) against the oracle's labels.
- learned proposal at its spec: the 20,000 x 20,000, 50-label atlas; a
  2 x 512 network trained on the plain method's final labels of 500,000
  training cells; the proposal form's first and final labels on 2,000 query
  cells against the oracle's.

Every case prints its name before touching the GPU."""

import dataclasses
import sys
import time

import numpy as np
import pytest

from helpers import assert_bits_equal

pytestmark = pytest.mark.gpu
NQ = 2000


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from cusingler_amd import engine
    e = engine.Engine(0)
    yield e
    e.close()


def _train_network(eng, ds, hidden, epochs):
    """The network trained on the plain method's final labels of ds's training
    cells (labelled on the GPU)."""
    from cusingler_amd import engine, synth
    tprep = engine.Prep(synth.training_set(ds), 16)
    eng.copyin_score(tprep.generate(0))
    eng.copyin(tprep.generate(1))
    del tprep
    eng.get_label(0.8, 0.05, "plain", outputs=False)
    ylab, _ = eng.finetune(0.8, 0.05, 0)
    eng.mlp_init(hidden, seed=1)
    losses = eng.mlp_train(ylab.astype(np.int32), epochs=epochs, batch=1024, lr=1e-3, seed=2)
    params = eng.mlp_params()
    eng.destroy()
    eng.destroy_score()
    return params, losses, int(ylab.shape[0])


def _oracle(ds):
    from oracle import oracle as O
    t0 = time.perf_counter()
    orc = O.Oracle(ds, 16)
    of, om, osc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    ofin, ornd = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    _say(f"  oracle on {ds.qry_height} cells: {time.perf_counter() - t0:.1f} s")
    return of, om, osc, ofin, ornd


def _proposal_labels(eng, c0, c1, params, hidden):
    eng.copyin_score(c0)
    eng.mlp_init(hidden, seed=1)
    eng.set_mlp_params(params)
    eng.set_proposal(4, 0.2)
    try:
        first, mask, _ = eng.get_label(0.8, 0.05, "exact_int8")
        cnt = eng.counters()
    finally:
        eng.set_proposal(0)
    eng.copyin(c1)
    final, _ = eng.finetune(0.8, 0.05, 0)
    eng.destroy()
    eng.destroy_score()
    return first, mask, final, cnt


def test_streaming_configuration(eng):
    from cusingler_amd import engine, synth
    cfg = dataclasses.replace(synth.CONFIGS["streaming"], n_qry=NQ, n_train=20000)
    _say("streaming: generate", cfg.n_ref, "x", cfg.n_common, "reference,", NQ, "query")
    ds = synth.generate(cfg)
    params, losses, ntrain = _train_network(eng, ds, 512, 6)
    _say("  network trained on", ntrain, "cells, losses", np.round(losses, 4).tolist())
    ds.train = None
    of, om, osc, ofin, ornd = _oracle(ds)
    prep = engine.Prep(ds, 16)
    c0, c1 = prep.generate(0), prep.generate(1)
    assert len(ds.codes) == 100 and ds.ref_height == 50000
    for path in ("plain", "exact_int8"):
        _say("streaming:", path, "G_s", c0.width, "G_f", c1.width)
        eng.copyin_score(c0)
        first, mask, scores = eng.get_label(0.8, 0.05, path)
        eng.copyin(c1)
        final, rounds = eng.finetune(0.8, 0.05, 0)
        eng.destroy()
        eng.destroy_score()
        if path == "plain":
            assert_bits_equal(scores, osc, "streaming plain scores")
        assert np.array_equal(first, of) and np.array_equal(mask, om), path
        assert np.array_equal(final, ofin) and np.array_equal(rounds, ornd), path
    _say("streaming: chosen form (exact INT8 + proposal)")
    first, mask, final, cnt = _proposal_labels(eng, c0, c1, params, 512)
    a1, a2 = float((first == of).mean()), float((final == ofin).mean())
    _say(f"  agreement with the oracle: first {a1:.5f} final {a2:.5f}; proposal-only cells "
         f"{cnt['prop_cells']}, scored pairs {cnt['prop_scored_pairs']} of {NQ * 100}")
    assert a1 >= 0.999 and a2 >= 0.999


def test_learned_proposal_at_spec(eng):
    from cusingler_amd import engine, synth
    cfg = dataclasses.replace(synth.CONFIGS["learned"], n_qry=NQ)
    assert cfg.n_train == 500000
    _say("learned: generate atlas with", cfg.n_train, "training cells")
    ds = synth.generate(cfg)
    params, losses, ntrain = _train_network(eng, ds, 512, 4)
    _say("  network 2 x 512 trained on", ntrain, "cells, losses", np.round(losses, 5).tolist())
    assert ntrain == 500000 and losses[-1] < 0.1 * losses[0]
    ds.train = None
    of, om, _, ofin, _ = _oracle(ds)
    prep = engine.Prep(ds, 16)
    c0, c1 = prep.generate(0), prep.generate(1)
    first, mask, final, cnt = _proposal_labels(eng, c0, c1, params, 512)
    a1, a2 = float((first == of).mean()), float((final == ofin).mean())
    frac = cnt["prop_scored_pairs"] / (NQ * len(ds.codes))
    _say(f"  agreement with the oracle: first {a1:.5f} final {a2:.5f}; pairs scored {frac:.3f}; "
         f"fallback margin {cnt['prop_fallback_margin']}, edge {cnt['prop_fallback_check']}")
    assert a1 >= 0.999 and a2 >= 0.999
    assert frac < 0.35
