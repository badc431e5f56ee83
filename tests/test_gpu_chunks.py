"""The score step over several query chunks (csg_set_query_chunk caps the
chunk): the chunk offsets of the plain, exact INT8 and learned-proposal paths
against the CPU oracle and against the one-chunk run, including the last
partial chunk, NaN cells and the fine-tune on the device mask that follows.

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
    yield e
    e.set_query_chunk(0)
    e.close()


def _data(name):
    from cusingler_amd import synth
    return {
        "small": lambda: synth.generate("small", n_qry=1000),
        # 330 query cells (a partial last chunk) with empty rows in several chunks
        "tiny_nan": lambda: synth.with_empty_rows(
            synth.tiny(n_ref_per_label=10, n_labels=3, n_genes=320, n_qry=330, seed=21),
            [3, 17], [5, 140, 301]),
    }[name]()


_ORACLE = {}


def _oracle(name):
    if name not in _ORACLE:
        from oracle import oracle as O
        ds = _data(name)
        orc = O.Oracle(ds, 8)
        of, om, osc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
        ofin, ornd = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
        _ORACLE[name] = (ds, of, om, osc, ofin, ornd)
    return _ORACLE[name]


def _run(eng, ds, path, cap):
    from cusingler_amd import engine
    prep = engine.Prep(ds, 8)
    eng.set_query_chunk(cap)
    try:
        eng.copyin_score(prep.generate(0))
        chunk = eng.set_query_chunk(cap)
        first, mask, sc = eng.get_label(0.8, 0.05, path)
        eng.copyin(prep.generate(1))        # the fine-tune reads the device mask
        final, rounds = eng.finetune(0.8, 0.05, 0)
    finally:
        eng.set_query_chunk(0)
    eng.destroy()
    eng.destroy_score()
    return chunk, first, mask, sc, final, rounds


@pytest.mark.parametrize("path", ["plain", "exact_int8"])
@pytest.mark.parametrize("name", ["small", "tiny_nan"])
def test_multi_chunk_matches_oracle(eng, name, path):
    ds, of, om, osc, ofin, ornd = _oracle(name)
    _say("multi-chunk", name, path, "one chunk")
    _, f1, m1, s1, fin1, r1 = _run(eng, ds, path, 0)
    for cap in (128, 256):
        _say("multi-chunk", name, path, "cap", cap)
        chunk, f, m, s, fin, r = _run(eng, ds, path, cap)
        assert chunk == cap and ds.qry_height > cap, (chunk, cap)
        assert np.array_equal(f, of), f"{(f != of).sum()} first labels differ"
        assert np.array_equal(m, om), "masks differ"
        if path == "plain":
            assert_bits_equal(s, osc, f"{name} scores (cap {cap})")
        assert_bits_equal(s, s1, f"{name} {path} scores vs one chunk (cap {cap})")
        assert np.array_equal(fin, ofin), f"{(fin != ofin).sum()} final labels differ"
        assert np.array_equal(r, ornd), "round counts differ"
    assert np.array_equal(f1, of) and np.array_equal(fin1, ofin)


def test_multi_chunk_proposal(eng):
    """The learned proposal's per-chunk work lists: with 128- and 384-row
    chunks, first labels, masks, scores and final labels equal the one-chunk
    proposal run bit for bit, and the labels equal the oracle's."""
    from cusingler_amd import engine
    ds, of, om, osc, ofin, ornd = _oracle("small")
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    _say("proposal multi-chunk: train")
    eng.copyin_score(c0)
    eng.mlp_init(64, seed=5)
    eng.mlp_train(ofin.astype(np.int32), epochs=8, batch=256, lr=2e-3, seed=1)
    params = eng.mlp_params()
    eng.destroy_score()
    runs = {}
    try:
        for cap in (0, 128, 384):
            _say("proposal multi-chunk cap", cap)
            eng.set_query_chunk(cap)
            eng.copyin_score(c0)
            eng.set_mlp_params(params)
            eng.set_proposal(4, 0.2)
            first, mask, sc = eng.get_label(0.8, 0.05, "exact_int8")
            cnt = eng.counters()
            eng.set_proposal(0)
            eng.copyin(c1)
            final, _ = eng.finetune(0.8, 0.05, 0)
            eng.destroy()
            eng.destroy_score()
            runs[cap] = (first, mask, sc, final, cnt)
    finally:
        eng.set_proposal(0)
        eng.set_query_chunk(0)
    f0, m0, s0, fin0, c_0 = runs[0]
    _say("proposal cells", c_0["prop_cells"], "of", ds.qry_height)
    assert c_0["prop_cells"] > 0
    for cap in (128, 384):
        f, m, s, fin, _ = runs[cap]
        assert np.array_equal(f, f0) and np.array_equal(m, m0), cap
        assert_bits_equal(s, s0, f"proposal scores (cap {cap})")
        assert np.array_equal(fin, fin0), cap
    assert float((f0 == of).mean()) >= 0.999 and float((fin0 == ofin).mean()) >= 0.999


def test_query_chunk_argument(eng):
    from cusingler_amd import engine
    with pytest.raises(engine.CsgError):
        eng.set_query_chunk(100)
    eng.set_query_chunk(0)
