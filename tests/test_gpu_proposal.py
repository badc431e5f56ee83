"""The learned label proposal on the GPU: the network's training kernels
(BF16 MFMA forward/backward, FP32 Adam) against the CPU oracle's gradient,
training on labels of the plain method, and the proposal-restricted exact
score step against the full paths (scores, labels, and what the trained
network saves compared with an untrained one)."""

import dataclasses
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
    e.close()


def _loaded(eng, name="small", n_qry=None):
    from cusingler_amd import engine, synth
    ds = synth.generate(name, n_qry=n_qry)
    prep = engine.Prep(ds, 8)
    c0 = prep.generate(0)
    eng.copyin_score(c0)
    return ds, c0


@pytest.mark.parametrize("hidden,n", [(64, 96), (128, 37)])
def test_gradient_matches_cpu_oracle(eng, hidden, n):
    from oracle import mlp_oracle as M
    _say("mlp gradient", hidden, n)
    _, c0 = _loaded(eng, "small", n_qry=300)
    eng.mlp_init(hidden, seed=11)
    dm = eng.mlp_dims()
    d = M.Dims(G=dm["G"], Kp=dm["Kp"], H=dm["H"], L=dm["L"], Lp=dm["Lp"])
    assert d.n_params == dm["n_params"]
    rng = np.random.default_rng(hidden + n)
    x = rng.uniform(-1, 1, (n, d.G)).astype(np.float32)
    y = rng.integers(0, d.L, n).astype(np.int32)
    w = eng.mlp_params()
    g, loss = eng.mlp_grad(x, y)
    lo, go = M.loss_and_grad(w.astype(np.float64), x.astype(np.float64), y, d, bf16=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    offs = np.cumsum([0] + d.sizes)
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    for b in range(6):
        gb, ob = g[offs[b]:offs[b + 1]], go[offs[b]:offs[b + 1]]
        rel = np.linalg.norm(gb - ob) / max(np.linalg.norm(ob), 1e-30)
        _say(f"  {names[b]}: |g|={np.linalg.norm(ob):.3e} rel err={rel:.2e}")
        assert rel < 3e-3, (names[b], rel)
    # against the unrounded FP64 network: the BF16 operands cost well under 10%
    _, g64 = M.loss_and_grad(w.astype(np.float64), x.astype(np.float64), y, d, bf16=False)
    assert np.linalg.norm(g - g64) < 0.1 * np.linalg.norm(g64)


def _features(c0, n):
    """Network features of the first n coded query rows: doubled centred
    ranks over the score gene set (oracle ranks) divided by G_s."""
    from oracle import oracle as O
    G = c0.width
    x = np.zeros((n, G), np.float32)
    for q in range(n):
        row = np.zeros(G, np.uint16)
        a, b = c0.qry_indptr[q], c0.qry_indptr[q + 1]
        row[c0.qry_indices[a:b]] = c0.qry_data[a:b]
        X = 2.0 * O.rank(row).astype(np.float64) - (G + 1)
        x[q] = X.astype(np.float32) * np.float32(1.0 / G)
    return x


def test_adam_step_moves_against_gradient(eng):
    """One training step on one batch: the loss it reports is the gradient
    call's loss on the same features, and Adam's first step moves every
    parameter with a clear gradient by lr against the gradient's sign, while
    the inert padding parameters do not move."""
    _say("mlp adam step")
    ds, c0 = _loaded(eng, "small", n_qry=256)
    eng.mlp_init(64, seed=3)
    w0 = eng.mlp_params()
    y = (np.arange(256) % eng.L).astype(np.int32)
    x = _features(c0, 256)
    g, loss = eng.mlp_grad(x, y)
    losses = eng.mlp_train(y, epochs=1, batch=256, lr=1e-3, seed=0)
    assert losses[0] == pytest.approx(loss, rel=1e-5)
    w1 = eng.mlp_params()
    moved = w1.astype(np.float64) - w0
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert big.sum() > 1000
    np.testing.assert_allclose(moved[big], -1e-3 * np.sign(g[big]), rtol=2e-2)
    assert np.all(moved[g == 0] == 0)


@pytest.fixture(scope="module")
def atlas_lite(eng):
    """A reduced atlas (50 labels, near-duplicate families) with a network
    trained on 30,000 training cells labelled by the plain method."""
    from cusingler_amd import engine, synth
    cfg = dataclasses.replace(synth.CONFIGS["atlas"], n_ref=5000, n_qry=6000, n_train=30000)
    _say("atlas-lite generate")
    ds = synth.generate(cfg)
    tds = synth.training_set(ds)
    tprep = engine.Prep(tds, 8)
    eng.copyin_score(tprep.generate(0))
    eng.get_label(0.8, 0.05, "plain", outputs=False)
    eng.copyin(tprep.generate(1))
    ylab, _ = eng.finetune(0.8, 0.05, 0)
    _say("atlas-lite train")
    eng.mlp_init(512, seed=1)
    losses = eng.mlp_train(ylab.astype(np.int32), epochs=6, batch=512, lr=1e-3, seed=2)
    _say("losses", losses)
    params = eng.mlp_params()
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    eng.copyin_score(c0)
    eng.set_proposal(0)
    first_p, mask_p, sc_p = eng.get_label(0.8, 0.05, "plain")
    eng.copyin(c1)
    final_p, _ = eng.finetune(0.8, 0.05, 0)
    first_e, mask_e, sc_e = eng.get_label(0.8, 0.05, "exact_int8")
    return dict(ds=ds, c0=c0, c1=c1, losses=losses, params=params,
                plain=(first_p, mask_p, sc_p, final_p), exact=(first_e, mask_e, sc_e))


def test_training_reduces_loss_and_learns_plain_labels(eng, atlas_lite):
    losses = atlas_lite["losses"]
    assert losses[-1] < 0.5 * losses[0], losses
    eng.set_mlp_params(atlas_lite["params"])
    top, margin = eng.mlp_predict(4)
    final_p = atlas_lite["plain"][3]
    top1 = float((top[:, 0] == final_p).mean())
    top4 = float((top == final_p[:, None]).any(axis=1).mean())
    _say(f"held-out top-1 {top1:.4f} top-4 {top4:.4f} margin>=0.2 {(margin >= 0.2).mean():.3f}")
    assert top1 > 0.8 and top4 > 0.98


def _run_proposal(eng, atlas_lite, k=4, margin=0.2):
    eng.set_proposal(k, margin)
    first, mask, sc = eng.get_label(0.8, 0.05, "exact_int8")
    cnt = eng.counters()
    eng.set_proposal(0)
    eng.copyin(atlas_lite["c1"])
    final, _ = eng.finetune(0.8, 0.05, 0)
    return first, mask, sc, final, cnt


def _ranks_X(data, indices, indptr, rows, G):
    """Doubled centred ranks X = 2 rank - (G + 1) of coded CSR rows (oracle ranks)."""
    from oracle import oracle as O
    out = np.zeros((len(rows), G), np.int64)
    for t, q in enumerate(rows):
        row = np.zeros(G, np.uint16)
        a, b = indptr[q], indptr[q + 1]
        row[indices[a:b]] = data[a:b]
        out[t] = (2.0 * O.rank(row).astype(np.float64)).astype(np.int64) - (G + 1)
    return out


def _exact_quantile(c0, q, l):
    """The exact path's quantile of (cell q, label l) before certification:
    exact integer dot products times the two rows' inverse roots 1/sqrt(sum
    X^2) (correctly rounded doubles), in double, rounded to float, then the
    reference's percentile of the label's scores."""
    from cusingler_amd import engine
    G = c0.width
    s0, n = int(c0.ctidx[2 * l]), int(c0.ctidx[2 * l + 1])
    xq = _ranks_X(c0.qry_data, c0.qry_indices, c0.qry_indptr, [q], G)
    xr = _ranks_X(c0.ref_data, c0.ref_indices, c0.ref_indptr, range(s0, s0 + n), G)
    dots = (xr @ xq[0]).astype(np.float64)
    ir_q = 1.0 / np.sqrt(float((xq[0] * xq[0]).sum()))
    ir_r = 1.0 / np.sqrt((xr * xr).sum(axis=1).astype(np.float64))
    scores = ((dots * ir_q) * ir_r).astype(np.float32)
    return engine.host_percentile(scores, 0.8)


def test_proposal_scores_and_labels(eng, atlas_lite):
    eng.set_mlp_params(atlas_lite["params"])
    first, mask, sc, final, cnt = _run_proposal(eng, atlas_lite)
    first_p, mask_p, sc_p, final_p = atlas_lite["plain"]
    _, _, sc_e = atlas_lite["exact"]
    Nq, L = sc.shape
    agree_first = float((first == first_p).mean())
    agree_final = float((final == final_p).mean())
    frac = cnt["prop_scored_pairs"] / (Nq * L)
    _say(f"agreement first {agree_first:.5f} final {agree_final:.5f}; pairs scored {frac:.3f}; "
         f"proposal-only {cnt['prop_cells']}, fallback margin {cnt['prop_fallback_margin']}, "
         f"fallback check {cnt['prop_fallback_check']}, widened {cnt['prop_widened']}")
    # every scored (cell, label) quantile is the full exact path's value, or
    # the plain value where the certification re-scored it in FP32 order;
    # where the full path re-scored a pair this path did not, it is the exact
    # path's own value before certification
    fin = np.isfinite(sc)
    u, ue, up = sc.view(np.uint32), sc_e.view(np.uint32), sc_p.view(np.uint32)
    odd = np.argwhere(fin & (u != ue) & (u != up))
    _say(f"scored pairs differing from both full paths: {len(odd)}")
    assert len(odd) <= 200
    for q, l in odd:
        assert ue[q, l] == up[q, l], (q, l)  # the full exact path re-scored it
        ex = np.float32(_exact_quantile(atlas_lite["c0"], int(q), int(l)))
        assert ex.view(np.uint32) == u[q, l], (q, l, sc[q, l], ex)
    assert np.all(np.isneginf(sc[~fin]) | np.isnan(sc[~fin]))
    # unscored labels are never in the mask
    assert np.all(mask[np.isneginf(sc)] == 0)
    assert cnt["prop_cells"] + cnt["prop_fallback_margin"] + cnt["prop_fallback_check"] == Nq
    assert agree_final >= 0.999 and agree_first >= 0.999


def test_proposal_limb_forms_bit_equal(eng, atlas_lite):
    """The grouped contraction in the three-product form and in the folded
    four-product form: identical labels, mask and quantiles."""
    eng.set_mlp_params(atlas_lite["params"])
    eng.copyin_score(atlas_lite["c0"])
    assert atlas_lite["c0"].width <= 1008
    out = {}
    for form in (1, 0):
        _say("proposal, limb form", form)
        eng.set_limb_form(form)
        try:
            eng.set_proposal(4, 0.2)
            out[form] = eng.get_label(0.8, 0.05, "exact_int8")
        finally:
            eng.set_proposal(0)
            eng.set_limb_form(0)
    for a, b in zip(out[0], out[1]):
        assert np.asarray(a).tobytes() == np.asarray(b).tobytes()


def test_trained_network_gains_over_untrained(eng, atlas_lite):
    """What the network buys: the trained proposal scores a fraction of the
    (cell, label) pairs at >= 99.9% agreement; an untrained network of the
    same shape proposes with small margins, so nearly every cell falls back
    to full scoring (same labels as the exact path, no saving)."""
    Nq, L = atlas_lite["exact"][2].shape
    final_p = atlas_lite["plain"][3]
    eng.set_mlp_params(atlas_lite["params"])
    *_, final_t, cnt_t = _run_proposal(eng, atlas_lite)
    eng.mlp_init(512, seed=99)
    *_, final_u, cnt_u = _run_proposal(eng, atlas_lite)
    eng.set_mlp_params(atlas_lite["params"])
    frac_t = cnt_t["prop_scored_pairs"] / (Nq * L)
    frac_u = cnt_u["prop_scored_pairs"] / (Nq * L)
    _say(f"pairs scored: trained {frac_t:.3f}, untrained {frac_u:.3f}")
    assert frac_t < 0.35
    assert frac_u > 0.9
    assert float((final_t == final_p).mean()) >= 0.999
    assert float((final_u == final_p).mean()) >= 0.999


def test_proposal_argument_errors(eng, atlas_lite):
    from cusingler_amd import engine
    eng.set_mlp_params(atlas_lite["params"])
    eng.set_proposal(4, 0.2)
    with pytest.raises(engine.CsgError):
        eng.get_label(0.8, 0.05, "plain")
    eng.set_proposal(0)
    with pytest.raises(engine.CsgError):
        eng.set_proposal(9)
    with pytest.raises(engine.CsgError):
        eng.set_proposal(4, 1.5)
    with pytest.raises(engine.CsgError):
        eng.mlp_grad(np.zeros((2, eng.mlp_dims()["G"]), np.float32), np.array([0, 99], np.int32))
