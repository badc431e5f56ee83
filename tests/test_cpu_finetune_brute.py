"""Pins the oracle's fine-tune (finetune_round / cufinetune) against an
independent FP64 brute force written from the reference's description of the
round (src/cusingler.cu:1138-1333), not from the oracle's code:

- k_t = round(500 * (2/3)^log2(k)) with C rounding (half away from zero);
- G_sub = union over ORDERED pairs (i, j), i != j, of the first min(len, k_t)
  genes of the pair's thre-space window, in ascending gene order;
- fewer than 20 genes -> {front};
- ranks by scipy.stats.rankdata (average ties), Spearman in FP64, the
  per-label quantile by np.percentile(method="linear");
- first minimum / last maximum; a NaN maximum -> {front}; keep a label when
  score > min and score >= max - thre; labels equal to the maximum are
  inserted at the front, the others appended; an empty result -> {front};
- cufinetune: candidates are the mask's labels in ascending order, rounds run
  while more than one is left and the round limit allows; the final label is
  the front of the last list (0 for an empty list).

Decisions within 1e-4 of a tie, of the minimum or of the threshold may flip
between FP32 and FP64; such a cell is not compared from that round on, as in
the score-step brute force.  Every planted variant of a rule must disagree
with the oracle somewhere, so that the check is known to bite.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.stats

from oracle import oracle as O

Q = float(np.float32(0.8))          # the reference's float quantile 80.0f / 100
THRE = float(np.float32(0.05))      # finetune_thre as a float
EPS = 1e-4


class Brute:
    """FP64 brute force of the fine-tune on one dataset (tables and coded
    matrices come from the oracle's preprocessing, which the reference's own
    io.cpp pins)."""

    def __init__(self, ds, mistake: str = ""):
        self.orc = O.Oracle(ds, 2)
        t = self.orc.tables()
        c = self.orc.coded(1)
        s = self.orc.sizes
        self.L, self.Gf = s["L"], s["Gf"]
        self.ctidx = t["ctidx"].reshape(self.L, 2)
        self.win = t["ctdidx1"].reshape(self.L, self.L, 2)
        self.diff = t["ctdiff1"]
        self.ref = self._dense(c["ref_data"], c["ref_indices"], c["ref_indptr"], s["Nr"])
        self.qry = self._dense(c["qry_data"], c["qry_indices"], c["qry_indptr"], s["Nq"])
        self.mistake = mistake

    def _dense(self, data, ind, ptr, n):
        out = np.zeros((n, self.Gf), np.float64)
        for r in range(n):
            out[r, ind[ptr[r]:ptr[r + 1]]] = data[ptr[r]:ptr[r + 1]]
        return out

    def k_t(self, k: int) -> int:
        base = 1.5 if self.mistake == "k_t_base" else 2 / 3
        x = 500.0 * base ** math.log2(k)
        return int(math.floor(x + 0.5))

    def genes(self, top):
        kt = self.k_t(len(top))
        uni = set()
        for i in top:
            for j in top:
                if i == j or (self.mistake == "unordered_pairs" and i > j):
                    continue
                pos, ln = self.win[i, j]
                uni.update(int(g) for g in self.diff[pos:pos + min(int(ln), kt)])
        return sorted(uni)

    @staticmethod
    def _spearman(x, Y):
        xc = x - x.mean()
        Yc = Y - Y.mean(axis=1, keepdims=True)
        den = np.sqrt((xc * xc).sum() * (Yc * Yc).sum(axis=1))
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(den > 0, (Yc @ xc) / np.where(den > 0, den, 1), np.nan)

    def round(self, q: int, top):
        """Returns (new list, FP64 scores or None, ambiguous)."""
        g = self.genes(top)
        min_genes = 10 if self.mistake == "min_genes_10" else 20
        if len(g) < min_genes:
            return [top[0]], None, False
        x = scipy.stats.rankdata(self.qry[q, g])
        P = []
        for lab in top:
            st, ln = self.ctidx[lab]
            R = np.apply_along_axis(scipy.stats.rankdata, 1, self.ref[st:st + ln][:, g])
            s = self._spearman(x, R)
            if len(s) <= 1:
                P.append(float(s[0]))
            elif np.isnan(s).any():
                P.append(float("nan"))
            else:
                P.append(float(np.percentile(s, 100 * Q, method="linear")))
        P = np.array(P)
        if np.isnan(P).all():
            return [top[0]], P, False
        if np.isnan(P).any():
            return None, P, True          # libstdc++ NaN ordering: not brute-forced
        mn, mx = P.min(), P.max()
        imin = int(np.argmin(P))                      # first minimum
        imax = len(P) - 1 - int(np.argmax(P[::-1]))   # last maximum
        th = mx - THRE
        amb = False
        for i, v in enumerate(P):
            if i != imin and v - mn < EPS:
                amb = True
            if i != imax and mx - v < EPS:
                amb = True
            if abs(v - th) < EPS:
                amb = True
        res = []
        for i, v in enumerate(P):
            low = v < mn if self.mistake == "keep_min" else v <= mn
            if low or v < th:
                continue
            if i == imax and self.mistake != "ascending":
                res.insert(0, top[i])
            else:
                res.append(top[i])
        if not res:
            res = [top[0]]
        return res, P, amb

    def finetune(self, q: int, mask_row, times: int = 0):
        """Returns (final, rounds, lists per round, ambiguous)."""
        top = [l for l in range(self.L) if mask_row[l] != 0]
        rounds, cnt, lists = 0, 0, []
        while len(top) > 1 and (times == 0 or cnt < times):
            if cnt >= self.L:
                # each true round drops its minimum, so a list shrinks within
                # L rounds; only a planted rule can keep it from shrinking
                return -1, rounds, lists, False
            cnt += 1
            new, P, amb = self.round(q, top)
            if amb:
                return None, None, lists, True
            lists.append((list(top), new, P))
            top = new
            rounds += 1
        return (top[0] if top else 0), rounds, lists, False


def _datasets():
    from cusingler_amd import synth
    return {
        "tiny3": synth.tiny(n_ref_per_label=8, n_labels=3, n_genes=320, n_qry=24, seed=3),
        "tiny6": synth.tiny(n_ref_per_label=7, n_labels=6, n_genes=400, n_qry=24, markers=8,
                            seed=9),
        "dirichlet5": synth.tiny(n_ref_per_label=9, n_labels=5, n_genes=360, n_qry=20,
                                 props="dirichlet", seed=5),
        "small_genes": synth.tiny(n_ref_per_label=8, n_labels=3, n_genes=320, n_qry=12,
                                  nonzero=0.1, markers=3, seed=2),
        # dense rows: pair windows of a few hundred genes, longer than k_t
        "dense10": synth.tiny(n_ref_per_label=5, n_labels=10, n_genes=1000, n_qry=8,
                              nonzero=0.6, markers=20, seed=13),
    }


_DS = None


def ds(name):
    global _DS
    if _DS is None:
        _DS = _datasets()
    return _DS[name]


def _masks(L, n, rng):
    """All labels, then random subsets of 2..L labels per cell."""
    out = [np.ones(L, np.float32)]
    for _ in range(n):
        k = int(rng.integers(2, L + 1))
        m = np.zeros(L, np.float32)
        m[rng.choice(L, size=k, replace=False)] = 1
        out.append(m)
    return out


def _compare(name, times, mistake="", n_masks=3, seed=0):
    """Returns (cells compared, cells that disagree)."""
    b = Brute(ds(name), mistake)
    orc = b.orc
    rng = np.random.default_rng(seed)
    Nq = orc.sizes["Nq"]
    checked, bad = 0, 0
    _, smask, _ = orc.score(0, Nq)
    for q in range(Nq):
        for m in [smask[q]] + _masks(b.L, n_masks, rng):
            fin, rnd, lists, amb = b.finetune(q, m, times)
            if amb:
                continue
            of, orr = orc.finetune(q, q + 1, m[None, :], Q, THRE, times)
            same = int(of[0]) == fin and int(orr[0]) == rnd
            for top, new, P in lists:
                o_new, o_sc = orc.finetune_round(q, top, Q, THRE)
                same = same and o_new == new
                if P is not None and not np.isnan(P).all():
                    same = same and np.allclose(o_sc, P, atol=2e-5)
            checked += 1
            bad += 0 if same else 1
    return checked, bad


ALL = ["tiny3", "tiny6", "dirichlet5", "small_genes", "dense10"]


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("times", [0, 1, 2])
def test_finetune_matches_fp64_brute_force(name, times):
    checked, bad = _compare(name, times)
    assert bad == 0, f"{bad} of {checked} cells differ from the FP64 brute force"
    assert checked >= 20


@pytest.mark.parametrize("mistake,times", [("ascending", 1), ("k_t_base", 0),
                                           ("unordered_pairs", 0), ("keep_min", 0),
                                           ("min_genes_10", 0)])
def test_planted_mistake_is_caught(mistake, times):
    """Each plausible slip in a rule makes the brute force disagree with the
    oracle on some dataset."""
    bad = 0
    for name in ALL:
        bad += _compare(name, times, mistake)[1]
        if bad:
            break
    assert bad > 0, mistake


def test_special_cases():
    b = Brute(ds("tiny6"))
    orc = b.orc
    L = b.L
    # one candidate: no round, final = that label
    for lab in range(L):
        m = np.zeros((1, L), np.float32)
        m[0, lab] = 1
        f, r = orc.finetune(0, 1, m)
        assert int(f[0]) == lab and int(r[0]) == 0
    # empty candidate list: final 0, no round
    f, r = orc.finetune(0, 1, np.zeros((1, L), np.float32))
    assert int(f[0]) == 0 and int(r[0]) == 0
    # fewer than 20 genes in the union: {front} after one round
    b2 = Brute(ds("small_genes"))
    o2 = b2.orc
    hits = 0
    for i in range(b2.L):
        for j in range(i + 1, b2.L):
            if len(b2.genes([i, j])) < 20:
                m = np.zeros((1, b2.L), np.float32)
                m[0, [i, j]] = 1
                f, r = o2.finetune(0, 1, m)
                assert int(f[0]) == i and int(r[0]) == 1
                assert o2.finetune_round(0, [j, i])[0] == [j]
                hits += 1
    assert hits > 0
    # all-NaN round (a query cell with no stored value): {front}
    from cusingler_amd import synth
    e = synth.with_empty_rows(ds("tiny6"), [], [4])
    be = Brute(e)
    top = list(range(be.L))
    new, P, amb = be.round(4, top)
    assert new == [0] and np.isnan(P).all() and not amb
    got, sc = be.orc.finetune_round(4, top)
    assert got == [0] and np.isnan(sc).all()
    f, r = be.orc.finetune(4, 5, np.ones((1, be.L), np.float32))
    assert int(f[0]) == 0 and int(r[0]) == 1


def test_k_t_values():
    """k_t against the values tabulated for the reference (C round on doubles)."""
    b = Brute.__new__(Brute)
    b.mistake = ""
    table = {2: 333, 3: 263, 4: 222, 5: 195, 8: 148, 10: 130, 12: 117, 14: 107, 20: 87,
             50: 51, 100: 34}
    for k, v in table.items():
        assert b.k_t(k) == v
