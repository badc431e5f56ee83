"""The MFMA score paths on the GPU: operand layouts, the exact INT8-limb path
(labels bit-exact to the reference order, scores within the certified width)
and the approximate FP16 / FP32-MFMA paths (error against the FP64 oracle)."""

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


@pytest.mark.parametrize("which,name", [(0, "int8"), (4, "int8_wide"), (5, "int8_x3"),
                                        (6, "int8_wide_three_product"), (3, "plain"),
                                        (1, "fp16"), (2, "fp32_mfma")])
def test_contraction_selftest(eng, which, name):
    _say("selftest", name)
    err = eng.selftest(which)
    if which in (0, 3, 4, 5, 6):
        assert err == 0.0, f"{name} contraction differs from the host ({err})"
    else:
        assert err < 1e-5, f"{name} relative error {err}"


def test_topk_kernels_bit_equal(eng):
    """The proposal's top-k in one thread per row (up to 64 labels) equals the
    wave-per-row kernel: labels and margin bits, with ties, -inf and NaN rows."""
    _say("selftest top-k")
    assert eng.selftest(7) == 0


@pytest.mark.parametrize("gs", [20, 64, 127, 128, 129, 255, 256, 257, 300, 1000])
def test_finetune_chunk_tree_bit_exact(eng, gs):
    _say("ft chunk tree", gs)
    assert eng.selftest_ft(gs) == 0


def _data(name):
    from cusingler_amd import synth
    return {
        "tiny": lambda: synth.tiny(),
        "tiny_dirichlet": lambda: synth.tiny(n_ref_per_label=9, n_labels=4, n_genes=400, n_qry=40,
                                             props="dirichlet", seed=5),
        "small": lambda: synth.generate("small"),
        "pancreas_2k": lambda: synth.generate("pancreas", n_qry=2000),
        "atlas_1k": lambda: synth.generate("atlas", n_qry=1000),
        "tiny_nan": lambda: synth.with_empty_rows(
            synth.tiny(n_ref_per_label=10, n_labels=3, n_genes=320, n_qry=30, seed=21),
            [3, 17], [5, 11]),
    }[name]()


@pytest.mark.parametrize("name", ["tiny", "tiny_dirichlet", "tiny_nan", "small", "pancreas_2k",
                                  "atlas_1k"])
def test_exact_path_labels_match_reference_order(eng, name):
    from cusingler_amd import engine
    from oracle import oracle as O
    ds = _data(name)
    prep = engine.Prep(ds, 8)
    c0 = prep.generate(0)
    _say("exact_int8", name, ds.qry_height, ds.ref_height, c0.width)
    eng.copyin_score(c0)
    first, mask, scores = eng.get_label(0.8, 0.05, "exact_int8")
    cnt = eng.counters()
    eng.destroy_score()
    orc = O.Oracle(ds, 8)
    of, om, osc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    assert np.array_equal(first, of), f"{(first != of).sum()} first labels differ"
    assert np.array_equal(mask, om), f"{(mask != om).sum()} mask entries differ"
    w = engine.certify_width(c0.width)
    assert np.array_equal(np.isnan(scores), np.isnan(osc)), "NaN quantiles differ"
    err = np.abs(scores.astype(np.float64) - osc.astype(np.float64))
    err = err[~np.isnan(err)] if np.isnan(err).any() else err
    assert err.size == 0 or np.max(err) <= w, f"score error {np.nanmax(err)} above the certified width {w}"
    # every exact-path quantile is the exact correlation's (normalised in
    # FP64: within one float ulp of the oracle's exact variant) or, where the
    # certification re-scored the pair, the reference's FP32 value
    _, _, oex = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_EXACT)
    ok = ~np.isnan(osc)
    ulp = np.spacing(np.abs(oex[ok]).astype(np.float32))
    near_exact = np.abs(scores[ok].astype(np.float64) - oex[ok]) <= ulp
    is_fp32 = scores[ok].view(np.uint32) == osc[ok].view(np.uint32)
    assert np.all(near_exact | is_fp32), f"{int((~(near_exact | is_fp32)).sum())} quantiles"
    # against the reference's FP32 quantiles: within 1e-5 relative, except
    # where the reference's own FP32 error (about 3.5e-6 absolute at 900
    # genes) is above 1e-5 of a small quantile
    rel = np.abs(scores[ok].astype(np.float64) - osc[ok]) / np.abs(osc[ok].astype(np.float64))
    big = np.abs(osc[ok]) >= 0.25
    assert not big.any() or rel[big].max() <= 1e-5, rel[big].max()
    print(f"[exact-rel] {name}: re-scored (FP32) quantiles {int((is_fp32 & ~near_exact).sum())}, "
          f"max rel vs FP32 reference {rel.max() if rel.size else 0:.3g}, at |P|>=0.25 "
          f"{rel[big].max() if big.any() else 0:.3g}; {(rel > 1e-5).sum()} of {rel.size} above "
          f"1e-5 (|P| <= {np.abs(osc[ok])[rel > 1e-5].max() if (rel > 1e-5).any() else 0:.3g})",
          flush=True)
    print(f"[exact] {name}: rechecked pairs {cnt['rechecked_pairs']}, uncertified cells "
          f"{cnt['uncertified_cells']} of {ds.qry_height}, max |dP| {np.max(err) if err.size else 0:.3g} "
          f"(width {w:.3g})", flush=True)


@pytest.mark.parametrize("path", ["fp16", "fp32_mfma"])
def test_approximate_paths_error(eng, path):
    from cusingler_amd import engine
    from oracle import oracle as O
    ds = _data("pancreas_2k")
    prep = engine.Prep(ds, 8)
    c0 = prep.generate(0)
    _say(path, "pancreas_2k")
    eng.copyin_score(c0)
    first, mask, scores = eng.get_label(0.8, 0.05, path)
    eng.destroy_score()
    orc = O.Oracle(ds, 8)
    of, om, osc = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    err = np.nanmax(np.abs(scores.astype(np.float64) - osc.astype(np.float64)))
    agree = (first == of).mean()
    print(f"[{path}] max |dP| vs reference order {err:.3g}, first-label agreement {agree:.5f}",
          flush=True)
    assert err < 1e-4
    assert agree > 0.99


@pytest.mark.parametrize("schedule", ["device", "host"])
@pytest.mark.parametrize("name", ["tiny", "tiny_dirichlet", "tiny_nan", "small", "pancreas_2k",
                                  "atlas_1k"])
def test_exact_finetune_matches_reference_order(eng, name, schedule):
    """The certified exact fine-tune gives the reference's final labels and
    round counts (cells it cannot certify are redone in FP32 order)."""
    from cusingler_amd import engine
    from oracle import oracle as O
    ds = _data(name)
    prep = engine.Prep(ds, 8)
    orc = O.Oracle(ds, 8)
    _, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    of, orounds = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    _say("exact finetune", name, schedule, ds.qry_height)
    eng.copyin(prep.generate(1), om)
    eng.set_finetune_path("exact_int8")
    eng.set_finetune_schedule(schedule)
    try:
        final, rounds = eng.finetune(0.8, 0.05, 0)
    finally:
        eng.set_finetune_schedule("device")
        eng.set_finetune_path("plain")
    cnt = eng.counters()
    eng.destroy()
    assert np.array_equal(final, of), f"{(final != of).sum()} final labels differ"
    assert np.array_equal(rounds, orounds)
    print(f"[exact-ft] {name}: redone cells {cnt['ft_redone_cells']} of {cnt['ft_items']} "
          f"(cell, round) items", flush=True)


def test_device_finetune_graph_replays_and_autotuner(eng):
    """The device-resident fine-tune replays its captured round graphs across
    rounds and calls: three calls in a row on one engine (with eager work and
    synchronisations between them), autotuned and untuned, all give the
    host-grouped schedule's labels and round counts."""
    from cusingler_amd import engine, synth
    ds = synth.generate("atlas", n_qry=8000, n_ref=5000)
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    _say("device fine-tune replays", ds.qry_height)
    eng.copyin_score(c0)
    eng.get_label(0.8, 0.05, "exact_int8", outputs=False)
    eng.copyin(c1)
    eng.set_finetune_path("exact_int8")
    try:
        eng.set_finetune_schedule("host")
        ref_final, ref_rounds = eng.finetune(0.8, 0.05, 0)
        eng.set_finetune_schedule("device")
        for autotune in (True, False):
            eng.finetune_tuning(autotune)
            for call in range(3):
                final, rounds = eng.finetune(0.8, 0.05, 0)
                assert np.array_equal(final, ref_final), (autotune, call)
                assert np.array_equal(rounds, ref_rounds), (autotune, call)
                # eager work between the replays
                eng.get_label(0.8, 0.05, "plain", outputs=True)
            t = eng.finetune_tuning()
            _say("tuning", t)
            assert t["tuned"] == autotune
        eng.finetune_tuning(True)
    finally:
        eng.set_finetune_schedule("device")
        eng.set_finetune_path("plain")
    eng.destroy()
    eng.destroy_score()


@pytest.mark.parametrize("name", ["pancreas_2k", "atlas_1k"])
def test_correlation_error_vs_fp64(eng, name):
    """Per-pair correlations of every path against the FP64 oracle (exact
    integer rank sums, FP64 normalisation): the plain path equals the
    reference's FP32 order bit for bit, the exact path stays within 1e-6, the
    approximate paths' errors are stated."""
    from cusingler_amd import engine
    from oracle import oracle as O
    ds = _data(name)
    prep = engine.Prep(ds, 8)
    c0 = prep.generate(0)
    eng.copyin_score(c0)
    nq = min(256, ds.qry_height)
    orc = O.Oracle(ds, 8)
    seq, f64 = orc.score_matrix(0, nq, O.VAR_SEQ, want64=True)
    errs = {}
    for path in ["plain", "exact_int8", "fp32_mfma", "fp16"]:
        s = eng.score_matrix(path, 0, nq).astype(np.float64)
        ok = ~np.isnan(f64)
        assert np.array_equal(np.isnan(s), ~ok), path
        errs[path] = float(np.max(np.abs(s[ok] - f64[ok])))
        if path == "plain":
            assert_bits_equal(s.astype(np.float32), seq, f"{name} plain scores")
    eng.destroy_score()
    _say(f"{name} (G_s={c0.width}, {nq} x {ds.ref_height} pairs) max |corr - FP64|: " +
         ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert errs["exact_int8"] <= 1e-6
    assert errs["plain"] <= 1e-5
    assert errs["fp32_mfma"] <= 1e-5
    assert errs["fp16"] <= 1e-4


@pytest.mark.parametrize("name", ["tiny_nan", "pancreas_2k", "atlas_1k"])
def test_exact_limb_forms_bit_equal(eng, name):
    """The exact contraction's three-product form (gene sets of up to 1,008
    genes) and the folded four-product form give the same score matrix bit for
    bit, and the same labels, mask and quantiles."""
    from cusingler_amd import engine
    ds = _data(name)
    c0 = engine.Prep(ds, 8).generate(0)
    assert c0.width <= 1008
    nq = min(512, ds.qry_height)
    out = {}
    for form in (1, 0):
        _say("exact limb form", form, name, "G_s", c0.width)
        eng.set_limb_form(form)
        try:
            eng.copyin_score(c0)
            s = eng.score_matrix("exact_int8", 0, nq)
            lab = eng.get_label(0.8, 0.05, "exact_int8")
            eng.destroy_score()
        finally:
            eng.set_limb_form(0)
        out[form] = (s, *lab)
    for a, b, what in zip(out[0], out[1], ["scores", "first", "mask", "quantiles"]):
        assert_bits_equal(np.asarray(a), np.asarray(b), f"{name} {what}")


def test_device_finetune_batches_oversized_rounds(eng):
    """A round whose work tables exceed the budget runs in group ranges: with
    a tiny budget every round is split, and labels and round counts still
    equal the host-grouped schedule's, in FP32 and exact form."""
    from cusingler_amd import engine, synth
    ds = synth.generate("small")
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    eng.copyin_score(c0)
    eng.get_label(0.8, 0.3, "plain", outputs=False)  # wide threshold: many groups
    eng.copyin(c1)
    try:
        for path in ("plain", "exact_int8"):
            eng.set_finetune_path(path)
            eng.set_finetune_schedule("host")
            ref_final, ref_rounds = eng.finetune(0.8, 0.3, 0)
            eng.set_finetune_schedule("device")
            before = eng.set_finetune_budget(1 << 16)
            final, rounds = eng.finetune(0.8, 0.3, 0)
            batches = eng.set_finetune_budget(8 << 30) - before
            _say("batched fine-tune", path, "range batches", batches)
            assert batches > 1
            assert np.array_equal(final, ref_final), path
            assert np.array_equal(rounds, ref_rounds), path
    finally:
        eng.set_finetune_budget(8 << 30)
        eng.set_finetune_schedule("device")
        eng.set_finetune_path("plain")
    eng.destroy()
    eng.destroy_score()


def test_device_finetune_fused_form(eng):
    """Groups of at most four cells take the fused FP32 form (ranks and
    correlations in one pass, no stored reference ranks): labels and round
    counts equal those with the fused form off and the host-grouped
    schedule's, in FP32 and exact form."""
    from cusingler_amd import engine, synth
    ds = synth.generate("small")
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    eng.copyin_score(c0)
    eng.get_label(0.8, 0.3, "plain", outputs=False)  # wide threshold: many small groups
    eng.copyin(c1)
    try:
        for path in ("plain", "exact_int8"):
            eng.set_finetune_path(path)
            eng.set_finetune_schedule("host")
            ref_final, ref_rounds = eng.finetune(0.8, 0.3, 0)
            eng.set_finetune_schedule("device")
            out = {}
            for fuse in (0, 1, 4):
                eng.set_finetune_fuse(fuse)
                final, rounds = eng.finetune(0.8, 0.3, 0)
                out[fuse] = (final, rounds, eng.counters()["ft_fused_groups"])  # (this call's)
                _say("fused fine-tune", path, "max cells", fuse, "fused groups", out[fuse][2])
                assert np.array_equal(final, ref_final), (path, fuse)
                assert np.array_equal(rounds, ref_rounds), (path, fuse)
            assert out[0][2] == 0 and out[4][2] >= out[1][2] > 0
    finally:
        eng.set_finetune_fuse(4)
        eng.set_finetune_schedule("device")
        eng.set_finetune_path("plain")
    eng.destroy()
    eng.destroy_score()


def _jittered(ds, seed):
    """`ds` with every stored value scaled by its own factor in (1, 1.25):
    almost every value of a row is distinct, so rows hold hundreds of value
    codes (16-bit codes)."""
    import dataclasses
    rng = np.random.default_rng(seed)

    def j(d):
        return (d * (1.0 + rng.integers(1, 1000, size=d.shape[0]) / 4096.0)).astype(np.float32)

    return dataclasses.replace(ds, ref_data=j(ds.ref_data), qry_data=j(ds.qry_data))


@pytest.mark.parametrize("case", ["long_gene_sets", "wide_codes"])
def test_device_finetune_kernel_variants(eng, case):
    """The device fine-tune's kernel variants give the host-grouped
    schedule's labels and round counts: fine-tune gene sets above 1024 genes
    (the register-resident kernels' chunk lengths 9 and 10) and more than 256
    value codes (16-bit codes: the LDS-table rank and fused kernels), in FP32
    and exact form, fused form on and off."""
    from cusingler_amd import engine, synth
    if case == "long_gene_sets":
        # G_f = 1120 (the device fine-tune's LDS plan takes G_f up to about 1250)
        ds = synth.tiny(n_ref_per_label=40, n_labels=10, n_genes=4000, n_qry=300, markers=6,
                        seed=31)
    else:
        ds = _jittered(synth.tiny(n_ref_per_label=40, n_labels=8, n_genes=1500, n_qry=300,
                                  markers=8, nonzero=0.5, seed=33), 5)
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    codes = max(int(c1.ref_data.max()), int(c1.qry_data.max()))
    _say("kernel variants", case, "G_f", c1.width, "codes", codes)
    if case == "long_gene_sets":
        assert c1.width > 1024
    else:
        assert codes > 255
    eng.copyin_score(c0)
    eng.get_label(0.8, 0.3, "plain", outputs=False)  # wide threshold: many labels per cell
    eng.copyin(c1)
    try:
        for path in ("plain", "exact_int8"):
            eng.set_finetune_path(path)
            eng.set_finetune_schedule("host")
            ref_final, ref_rounds = eng.finetune(0.8, 0.3, 0)
            eng.set_finetune_schedule("device")
            for fuse in (0, 4):
                eng.set_finetune_fuse(fuse)
                final, rounds = eng.finetune(0.8, 0.3, 0)
                _say("kernel variants", case, path, "fuse", fuse)
                assert np.array_equal(final, ref_final), (case, path, fuse)
                assert np.array_equal(rounds, ref_rounds), (case, path, fuse)
    finally:
        eng.set_finetune_fuse(4)
        eng.set_finetune_schedule("device")
        eng.set_finetune_path("plain")
    eng.destroy()
    eng.destroy_score()


def test_phase_timing_modes(eng):
    """csg_set_phase_timing: mode 0 records no phase, mode 1 only the score
    step's correlation, mode 2 every phase; other modes are rejected.  The
    labels do not depend on the mode."""
    from cusingler_amd import engine, synth
    ds = synth.generate("small")
    prep = engine.Prep(ds, 8)
    c0 = prep.generate(0)
    eng.copyin_score(c0)
    labels = {}
    try:
        for mode in (0, 1, 2):
            eng.set_phase_timing(mode)
            eng.phase_times(reset=True)
            first, _, _ = eng.get_label(0.8, 0.05, "exact_int8", outputs=True)
            ph = eng.phase_times(reset=True)
            labels[mode] = first
            _say("phase timing mode", mode, {k: v[1] for k, v in ph.items()})
            timed = {k for k, v in ph.items() if v[1] > 0}
            if mode == 0:
                assert not timed
            elif mode == 1:
                assert timed == {"corr"}
            else:
                assert {"rank", "corr", "select", "label"} <= timed
        for bad in (-1, 4):
            with pytest.raises(engine.CsgError):
                eng.set_phase_timing(bad)
    finally:
        eng.set_phase_timing(2)
    assert np.array_equal(labels[0], labels[1]) and np.array_equal(labels[0], labels[2])
    eng.destroy_score()


def test_finetune_family_timing_mode(eng):
    """Phase mode 3 launches the device fine-tune's round work eagerly and
    times its kernel families (rank, correlation, select+filter); the labels
    and round counts equal the graph replays', and the work counter (cells x
    reference rows x |G_sub| over all rounds) is filled."""
    from cusingler_amd import engine, synth
    ds = synth.generate("small")
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    eng.copyin_score(c0)
    eng.get_label(0.8, 0.3, "plain", outputs=False)  # wide threshold: several rounds
    eng.copyin(c1)
    try:
        res = {}
        for mode in (2, 3):
            _say("fine-tune family timing, mode", mode)
            eng.set_phase_timing(mode)
            eng.phase_times(reset=True)
            final, rounds = eng.finetune(0.8, 0.3, 0)
            ph = eng.phase_times(reset=True)
            # (the fine-tune's counters cover its last call)
            res[mode] = (final, rounds, ph, eng.counters()["ft_pair_genes"])
    finally:
        eng.set_phase_timing(2)
    eng.destroy()
    eng.destroy_score()
    assert np.array_equal(res[2][0], res[3][0]) and np.array_equal(res[2][1], res[3][1])
    ph3 = res[3][2]
    _say("families (ms):", {k: round(ph3[k][0], 3) for k in ("ft_rank", "ft_corr", "ft_select")},
         "pair genes", res[3][3])
    assert ph3["ft_rank"][1] > 0 and ph3["ft_corr"][1] > 0 and ph3["ft_select"][1] > 0
    assert res[2][2]["ft_corr"][1] == 0
    assert res[2][3] == res[3][3] > 0
