"""Pins the CPU oracle against what mathematics and independent libraries fix:
scipy's average ranks, NumPy FP64 correlations and percentiles, a brute-force
FP64 restatement of the label rules on tiny inputs, and the golden fixtures
(whose preprocessing half comes from the reference's own src/io.cpp)."""

import glob
import os

import numpy as np
import pytest
import scipy.stats

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RNG = np.random.default_rng(20261203)


def test_rank_matches_scipy_average_ranks():
    for n in [1, 2, 7, 64, 300, 1000]:
        for hi in [1, 3, 20, 1000]:
            codes = RNG.integers(0, hi, size=n).astype(np.uint16)
            assert np.array_equal(O.rank(codes), scipy.stats.rankdata(codes, "average"))


@pytest.mark.parametrize("G", [5, 130, 901, 3000])
def test_correlation_variants_against_fp64(G):
    for _ in range(20):
        q = scipy.stats.rankdata(RNG.integers(0, 12, size=G)).astype(np.float32)
        r = scipy.stats.rankdata(RNG.integers(0, 12, size=G)).astype(np.float32)
        ref = np.corrcoef(q.astype(np.float64), r.astype(np.float64))[0, 1]
        for var, tol in [(O.VAR_SEQ, 2e-5 * max(1, G / 1000)), (O.VAR_TREE, 2e-6),
                         (O.VAR_EXACT, 1e-7)]:
            f, d = O.corr(q, r, var)
            assert abs(d - ref) < 1e-12
            assert abs(f - ref) < tol, (var, f, ref)


def test_correlation_special_cases():
    q = scipy.stats.rankdata(RNG.integers(0, 9, size=200)).astype(np.float32)
    for var in (O.VAR_SEQ, O.VAR_TREE, O.VAR_EXACT):
        assert O.corr(q, q, var)[0] == pytest.approx(1.0, abs=1e-6)
        assert O.corr(q, (201 - q).astype(np.float32), var)[0] == pytest.approx(-1.0, abs=1e-6)
        const = np.full(200, 100.5, np.float32)  # all-tied row: zero variance
        assert np.isnan(O.corr(q, const, var)[0])


def test_percentile_matches_numpy_linear():
    for n in [2, 3, 10, 100, 399, 400, 1000]:
        for _ in range(30):
            a = RNG.random(n).astype(np.float32)
            got = O.percentile(a, np.float32(0.8))
            want = np.percentile(a.astype(np.float64), 80.0000011920929, method="linear")
            assert abs(got - want) <= 4 * np.spacing(np.float32(want)), (n, got, want)
    a = RNG.random(50).astype(np.float32)
    assert O.percentile(a[:1], 0.8) == a[0]            # len 1: front
    assert O.percentile(a, 0.0) == a.min()             # p <= 0: smallest
    assert O.percentile(a, 1.0) == a.max()             # p >= last index: largest


def _brute_labels(ds, qrows, quantile=0.8, thre=0.05):
    """FP64 restatement of get_label on tiny inputs: ranks by scipy, Spearman by
    np.corrcoef, quantile by np.percentile, last maximum, mask by threshold."""
    orc = O.Oracle(ds, 2)
    t = orc.tables()
    c0 = orc.coded(0)
    G = len(t["common_genes"])

    def dense(data, ind, ptr, row):
        out = np.zeros(G)
        out[ind[ptr[row]:ptr[row + 1]]] = data[ptr[row]:ptr[row + 1]]
        return out
    Nr = orc.sizes["Nr"]
    R = np.array([scipy.stats.rankdata(dense(c0["ref_data"], c0["ref_indices"], c0["ref_indptr"], i))
                  for i in range(Nr)])
    L = orc.sizes["L"]
    ctidx = t["ctidx"]
    res = []
    for q in qrows:
        x = scipy.stats.rankdata(dense(c0["qry_data"], c0["qry_indices"], c0["qry_indptr"], q))
        s = np.array([np.corrcoef(x, R[i])[0, 1] for i in range(Nr)])
        P = np.array([np.percentile(s[ctidx[2 * l]:ctidx[2 * l] + ctidx[2 * l + 1]], 80,
                                    method="linear") for l in range(L)])
        first = L - 1 - int(np.argmax(P[::-1]))
        res.append((first, P >= P.max() - thre, P))
    return res


@pytest.mark.parametrize("seed", [3, 9, 27])
def test_score_step_against_fp64_brute_force(seed):
    from cusingler_amd import synth
    ds = synth.tiny(n_ref_per_label=8, n_labels=3, n_genes=320, n_qry=12, seed=seed)
    orc = O.Oracle(ds, 2)
    first, mask, scores = orc.score(0, ds.qry_height)
    brute = _brute_labels(ds, range(ds.qry_height))
    checked = 0
    for q, (bf, bm, bp) in enumerate(brute):
        assert np.allclose(scores[q], bp, atol=2e-5)
        srt = np.sort(bp)
        margin = srt[-1] - srt[-2]
        near_thre = np.min(np.abs(bp - (bp.max() - 0.05)))
        if margin > 1e-4:
            assert first[q] == bf
            checked += 1
        if near_thre > 1e-4:
            assert np.array_equal(mask[q] != 0, bm)
    assert checked > 0


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))))
def test_oracle_matches_golden_fixture(path):
    g = np.load(path)
    from cusingler_amd.synth import Dataset
    ds = Dataset(cfg=None, ref_height=int(g["ref_height"]), ref_width=int(g["ref_width"]),
                 ref_data=g["ref_data"], ref_indices=g["ref_indices"], ref_indptr=g["ref_indptr"],
                 qry_height=int(g["qry_height"]), qry_width=int(g["qry_width"]),
                 qry_data=g["qry_data"], qry_indices=g["qry_indices"], qry_indptr=g["qry_indptr"],
                 codes=list(g["codes"]), celltype=g["celltype"], cellnames=list(g["cellnames"]),
                 ref_genes=list(g["ref_genes"]), qry_genes=list(g["qry_genes"]),
                 qry_truth=np.zeros(0, np.int32))
    orc = O.Oracle(ds, 2)
    # preprocessing against the reference's own src/io.cpp outputs
    t = orc.tables()
    for ours, theirs in [("common_genes", "ref_common_genes"), ("thre_genes", "ref_thre_genes"),
                         ("ctdiff1", "ref_ctdiff1"), ("ctdidx1", "ref_ctdidx1"),
                         ("train_values", "ref_train_values"), ("train_idxs", "ref_train_idxs")]:
        assert np.array_equal(t[ours], g[theirs]), ours
    for step in (0, 1):
        c = orc.coded(step)
        for k, v in c.items():
            assert np.array_equal(v, g[f"step{step}_{k}"]), (step, k)
    # score step and fine-tune against the fixture (regression)
    first, mask, scores = orc.score(0, ds.qry_height)
    assert np.array_equal(first, g["first"])
    assert np.array_equal(mask, g["mask"])
    assert np.array_equal(scores.view(np.uint32), g["scores"].view(np.uint32))
    final, rounds = orc.finetune(0, ds.qry_height, mask)
    assert np.array_equal(final, g["final"])
    assert np.array_equal(rounds, g["rounds"])


def test_planted_mistake_is_caught():
    """A percentile with a wrong interpolation end (a plausible slip) fails the
    NumPy check used above."""
    a = np.random.default_rng(5).random(100).astype(np.float32)
    srt = np.sort(a)
    wrong = srt[80]  # position 0.8 * 99 = 79.2 read without interpolation
    want = np.percentile(a.astype(np.float64), 80.0000011920929, method="linear")
    assert abs(O.percentile(a, 0.8) - want) < 1e-6
    assert abs(wrong - want) > 1e-4
