"""Seeded synthetic reference/query atlases for the SingleR annotation path.

Each dataset is a reference CSR matrix with per-cell label codes plus a query
CSR matrix, both holding log2(1 + count) as float32 with only counts >= 1
stored, and gene-name lists for both sides.  Counts are NegativeBinomial with
size 5 (dispersion 0.2); the mean of gene g in cell c is

    mu = s * base[g] * fold[label(c), g] * lib[c]

with base ~ LogNormal(0, 1), lib ~ LogNormal(8, 0.3) normalised to mean 1, and
s calibrated so that the expected nonzero fraction hits the configured target.
Labels up-regulate marker genes drawn from the 5% of genes with the highest
baseline (markers are well-expressed genes, so they are detected at 4-8%
density); near-duplicate labels share a fraction of their markers (families /
parent-child hierarchies), which is what makes the fine-tune step see several
candidate labels per cell: on the atlas the candidate sets are Poisson-skewed
over 1-14 labels (median 2), on the streaming hierarchy they are the ~10
children of one parent.

The per-entry sampling runs in ``libcsg_synth.so`` (OpenMP C++), keyed by
(seed, stream, row, gene) so the data does not depend on the thread count.
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import Dict, List, Optional

import numpy as np

from . import _lib

NB_SIZE = 5.0  # dispersion 0.2


@dataclasses.dataclass
class SynthConfig:
    name: str
    n_ref: int
    n_qry: int
    n_common: int            # genes shared by both sides
    n_ref_only: int          # genes only in the reference
    n_qry_only: int          # genes only in the query
    n_labels: int
    markers: int             # markers per label
    fold_lo: float
    fold_hi: float
    nonzero: float           # target nonzero fraction
    props: str = "uniform"   # "uniform" | "dirichlet"
    dirichlet_alpha: float = 0.5
    min_per_label: int = 2
    family: str = "none"     # "none" | "poisson" | "hierarchy"
    share: float = 0.0       # fraction of markers shared inside a family
    family_lambda: float = 2.0
    n_parents: int = 0
    seed: int = 1234
    n_train: int = 0         # extra query-like cells (learned proposal training set)
    marker_pool: float = 0.05  # markers are drawn from this top fraction of genes by baseline
    base_sigma: float = 1.0    # baseline expression ~ LogNormal(0, base_sigma)
    label_sigma: float = 0.0   # every gene's per-label fold ~ LogNormal(0, label_sigma) (0: none)


CONFIGS: Dict[str, SynthConfig] = {
    # 1,000 ref cells x 2,000 genes (10 labels x 100 cells), 2,000 query cells,
    # 20 markers per label at 4x, ~8% nonzero, genes fully shared.
    "small": SynthConfig("small", 1000, 2000, 2000, 0, 0, 10, 20, 4.0, 4.0, 0.08),
    # 2,000 ref x 15,000 genes, 14 labels Dirichlet(0.5); 8,000 query x 14,000
    # genes with 12,000 shared; 50 markers per label at 2-8x; ~6% nonzero.
    "pancreas": SynthConfig("pancreas", 2000, 8000, 12000, 3000, 2000, 14, 50, 2.0, 8.0, 0.06,
                            props="dirichlet"),
    # 20,000 ref x 20,000 genes, 50 labels, 100 markers; 200,000 query cells at
    # 5% nonzero; near-duplicate label families sharing 60% of markers.
    "atlas": SynthConfig("atlas", 20000, 200000, 20000, 0, 0, 50, 100, 2.0, 8.0, 0.05,
                         family="poisson", share=0.6),
    # The atlas plus 500,000 query-like training cells for the learned proposal.
    "learned": SynthConfig("learned", 20000, 200000, 20000, 0, 0, 50, 100, 2.0, 8.0, 0.05,
                           family="poisson", share=0.6, n_train=500000),
    # The learned-proposal atlas with every gene's level varying between labels
    # (fold ~ LogNormal(0, 2)): label pairs differ in thousands of genes, so
    # G_s = 2,097 (over the folded-limb kernel's 1,920) and G_f = 8,183.
    "atlas_wide": SynthConfig("atlas_wide", 20000, 200000, 20000, 0, 0, 50, 100, 2.0, 8.0, 0.05,
                              family="poisson", share=0.6, n_train=500000, label_sigma=2.0),
    # 50,000 ref x 30,000 genes, 100 labels = 10 parents x 10 children sharing
    # 70% of markers; 1,000,000 query cells at 4% nonzero.
    "streaming": SynthConfig("streaming", 50000, 1000000, 30000, 0, 0, 100, 100, 2.0, 8.0, 0.04,
                             family="hierarchy", share=0.7, n_parents=10),
}


@dataclasses.dataclass
class Dataset:
    cfg: SynthConfig
    ref_height: int
    ref_width: int
    ref_data: np.ndarray      # float32 [nnz]
    ref_indices: np.ndarray   # int32 [nnz]
    ref_indptr: np.ndarray    # int64 [ref_height + 1]
    qry_height: int
    qry_width: int
    qry_data: np.ndarray
    qry_indices: np.ndarray
    qry_indptr: np.ndarray
    codes: List[str]          # label names in code order
    celltype: np.ndarray      # int32 per ref cell
    cellnames: List[str]
    ref_genes: List[str]
    qry_genes: List[str]
    qry_truth: np.ndarray     # int32 generating label of each query cell
    train: Optional["Dataset"] = None


def _synth_lib():
    lib = _lib.load("libcsg_synth.so")
    lib.csg_synth_sample.restype = ctypes.c_int64
    lib.csg_synth_sample.argtypes = [
        ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
        ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
    lib.csg_synth_free.argtypes = [ctypes.c_void_p]
    lib.csg_synth_nonzero_fraction.restype = ctypes.c_double
    lib.csg_synth_nonzero_fraction.argtypes = [
        ctypes.c_double, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_double]
    return lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _sample(lib, seed, stream, mu, labels, libf, force_min_nnz):
    n = labels.shape[0]
    g = mu.shape[1]
    indptr = np.zeros(n + 1, dtype=np.int64)
    pd = ctypes.c_void_p()
    pi = ctypes.c_void_p()
    nnz = lib.csg_synth_sample(seed, stream, n, g, _ptr(mu), _ptr(labels), _ptr(libf), NB_SIZE,
                               force_min_nnz, _ptr(indptr), ctypes.byref(pd), ctypes.byref(pi))
    data = np.empty(nnz, dtype=np.float32)
    idx = np.empty(nnz, dtype=np.int32)
    if nnz:
        ctypes.memmove(_ptr(data), pd.value, nnz * 4)
        ctypes.memmove(_ptr(idx), pi.value, nnz * 4)
    lib.csg_synth_free(pd)
    lib.csg_synth_free(pi)
    return data, idx, indptr


def _label_families(cfg: SynthConfig, rng: np.random.Generator) -> List[List[int]]:
    L = cfg.n_labels
    if cfg.family == "none":
        return [[i] for i in range(L)]
    if cfg.family == "hierarchy":
        per = L // cfg.n_parents
        return [list(range(p * per, (p + 1) * per)) for p in range(cfg.n_parents)]
    fams, start = [], 0
    while start < L:
        size = int(min(12, 1 + rng.poisson(cfg.family_lambda)))
        size = min(size, L - start)
        fams.append(list(range(start, start + size)))
        start += size
    return fams


def _fold_matrix(cfg: SynthConfig, rng: np.random.Generator, common_pos: np.ndarray,
                 n_universe: int, base: np.ndarray) -> np.ndarray:
    """Per-label fold over the gene universe; markers are drawn from shared genes."""
    L = cfg.n_labels
    fold = np.ones((L, n_universe), dtype=np.float32)
    fams = _label_families(cfg, rng)
    n_shared = int(round(cfg.markers * cfg.share)) if cfg.family != "none" else 0
    n_priv = cfg.markers - n_shared
    # candidates exclude the first 32 shared genes so the low-index genes stay
    # un-marked (they fill the zero-difference slots of the gene sets)
    pool = common_pos[32:] if common_pos.shape[0] > 64 + cfg.markers else common_pos
    if cfg.marker_pool < 1.0:
        # markers are genes that are well expressed at baseline (as real markers are)
        n_keep = max(4 * cfg.markers, int(round(pool.shape[0] * cfg.marker_pool)))
        order = np.argsort(-base[pool], kind="stable")
        pool = np.sort(pool[order[:n_keep]])
    for fam in fams:
        shared = rng.choice(pool, size=n_shared, replace=False) if n_shared else np.zeros(0, np.int64)
        shared_fold = rng.uniform(cfg.fold_lo, cfg.fold_hi, size=n_shared).astype(np.float32)
        for lab in fam:
            fold[lab, shared] = shared_fold
            priv = rng.choice(pool, size=n_priv, replace=False)
            fold[lab, priv] = rng.uniform(cfg.fold_lo, cfg.fold_hi, size=n_priv).astype(np.float32)
    return fold


def _proportions(cfg: SynthConfig, rng: np.random.Generator, n: int) -> np.ndarray:
    L = cfg.n_labels
    if cfg.props == "dirichlet":
        p = rng.dirichlet(np.full(L, cfg.dirichlet_alpha))
        counts = np.maximum(cfg.min_per_label, np.floor(p * n).astype(np.int64))
        while counts.sum() > n:
            i = int(np.argmax(counts))
            counts[i] -= 1
        while counts.sum() < n:
            counts[int(np.argmax(p))] += 1
    else:
        counts = np.full(L, n // L, dtype=np.int64)
        counts[: n - counts.sum()] += 1
    return counts


def _lib_factors(rng: np.random.Generator, n: int) -> np.ndarray:
    lib = rng.lognormal(8.0, 0.3, size=n)
    return (lib / lib.mean()).astype(np.float32)


def generate(cfg: SynthConfig | str, n_qry: Optional[int] = None, n_ref: Optional[int] = None,
             seed: Optional[int] = None) -> Dataset:
    """Generate one dataset; ``n_qry`` / ``n_ref`` override the configured sizes."""
    if isinstance(cfg, str):
        cfg = CONFIGS[cfg]
    if n_qry is not None or n_ref is not None or seed is not None:
        cfg = dataclasses.replace(cfg, n_qry=n_qry if n_qry is not None else cfg.n_qry,
                                  n_ref=n_ref if n_ref is not None else cfg.n_ref,
                                  seed=seed if seed is not None else cfg.seed)
    lib = _synth_lib()
    rng = np.random.default_rng(cfg.seed)
    n_uni = cfg.n_common + cfg.n_ref_only + cfg.n_qry_only
    # universe layout: [common | ref-only | qry-only]
    common = np.arange(cfg.n_common)
    ref_only = np.arange(cfg.n_common, cfg.n_common + cfg.n_ref_only)
    qry_only = np.arange(cfg.n_common + cfg.n_ref_only, n_uni)
    names = np.array([f"G{i:06d}" for i in range(n_uni)])

    def interleave(extra):
        # common genes keep their relative order (the entry point pairs the
        # shared genes of both sides by position)
        if extra.shape[0] == 0:
            return common.copy()
        slots = np.sort(rng.choice(cfg.n_common + extra.shape[0], size=extra.shape[0], replace=False))
        out = np.empty(cfg.n_common + extra.shape[0], dtype=np.int64)
        mask = np.zeros(out.shape[0], dtype=bool)
        mask[slots] = True
        out[mask] = rng.permutation(extra)
        out[~mask] = common
        return out

    ref_order = interleave(ref_only)
    qry_order = interleave(qry_only)

    base = rng.lognormal(0.0, cfg.base_sigma, size=n_uni).astype(np.float32)
    fold = _fold_matrix(cfg, rng, common, n_uni, base)
    if cfg.label_sigma > 0:
        # label programmes beyond the markers: every gene's mean differs a
        # little between labels, so label pairs differ in many genes
        fold *= rng.lognormal(0.0, cfg.label_sigma, size=fold.shape).astype(np.float32)
    mu_uni = base[None, :] * fold  # [L, universe]

    # reference labels (grouped sizes), shuffled cell order
    counts = _proportions(cfg, rng, cfg.n_ref)
    ref_labels = np.repeat(np.arange(cfg.n_labels, dtype=np.int32), counts)
    ref_labels = ref_labels[rng.permutation(ref_labels.shape[0])].astype(np.int32)
    qry_labels = rng.integers(0, cfg.n_labels, size=cfg.n_qry).astype(np.int32)
    ref_lib = _lib_factors(rng, cfg.n_ref)
    qry_lib = _lib_factors(rng, cfg.n_qry)

    # calibrate the global scale on a sample of cells
    cal_n = min(256, cfg.n_ref)
    cal_lab = ref_labels[:cal_n].copy()
    cal_lib = ref_lib[:cal_n].copy()
    mu_ref = np.ascontiguousarray(mu_uni[:, ref_order], dtype=np.float32)
    lo, hi = 1e-6, 1e3
    for _ in range(60):
        mid = math.sqrt(lo * hi)
        f = lib.csg_synth_nonzero_fraction(mid, cal_n, mu_ref.shape[1], _ptr(mu_ref), _ptr(cal_lab),
                                           _ptr(cal_lib), NB_SIZE)
        if f < cfg.nonzero:
            lo = mid
        else:
            hi = mid
    s = math.sqrt(lo * hi)
    mu_ref = (mu_ref * np.float32(s)).astype(np.float32)
    mu_qry = np.ascontiguousarray(mu_uni[:, qry_order] * np.float32(s), dtype=np.float32)

    rd, ri, rp = _sample(lib, cfg.seed, 1, mu_ref, ref_labels, ref_lib, 1)
    qd, qi, qp = _sample(lib, cfg.seed, 2, mu_qry, qry_labels, qry_lib, 1)

    ds = Dataset(cfg=cfg, ref_height=cfg.n_ref, ref_width=int(ref_order.shape[0]),
                 ref_data=rd, ref_indices=ri, ref_indptr=rp,
                 qry_height=cfg.n_qry, qry_width=int(qry_order.shape[0]),
                 qry_data=qd, qry_indices=qi, qry_indptr=qp,
                 codes=[f"type_{i:03d}" for i in range(cfg.n_labels)],
                 celltype=ref_labels,
                 cellnames=[f"cell_{i:07d}" for i in range(cfg.n_qry)],
                 ref_genes=list(names[ref_order]), qry_genes=list(names[qry_order]),
                 qry_truth=qry_labels)
    if cfg.n_train:
        tl = rng.integers(0, cfg.n_labels, size=cfg.n_train).astype(np.int32)
        tlib = _lib_factors(rng, cfg.n_train)
        td, ti, tp = _sample(lib, cfg.seed, 3, mu_qry, tl, tlib, 1)
        ds.train = Dataset(cfg=cfg, ref_height=0, ref_width=0, ref_data=np.zeros(0, np.float32),
                           ref_indices=np.zeros(0, np.int32), ref_indptr=np.zeros(1, np.int64),
                           qry_height=cfg.n_train, qry_width=ds.qry_width, qry_data=td,
                           qry_indices=ti, qry_indptr=tp, codes=ds.codes, celltype=ref_labels,
                           cellnames=[], ref_genes=ds.ref_genes, qry_genes=ds.qry_genes,
                           qry_truth=tl)
    return ds


def training_set(ds: Dataset) -> Dataset:
    """The reference of `ds` with its training cells (cfg.n_train) as the query
    set: the input on which the plain method labels the proposal network's
    training data."""
    t = ds.train
    if t is None:
        raise ValueError("dataset has no training cells (n_train = 0)")
    return dataclasses.replace(ds, qry_height=t.qry_height, qry_data=t.qry_data,
                               qry_indices=t.qry_indices, qry_indptr=t.qry_indptr,
                               cellnames=[f"train_{i:07d}" for i in range(t.qry_height)],
                               qry_truth=t.qry_truth, train=None)


def tiny(n_ref_per_label=6, n_labels=3, n_genes=300, n_qry=10, nonzero=0.3, markers=6,
         seed=7, n_ref_only=0, n_qry_only=0, props="uniform") -> Dataset:
    """Small dataset for brute-force checks."""
    n_common = n_genes
    cfg = SynthConfig("tiny", n_ref_per_label * n_labels, n_qry, n_common, n_ref_only, n_qry_only,
                      n_labels, markers, 3.0, 6.0, nonzero, props=props, seed=seed)
    return generate(cfg)


def with_empty_rows(ds: Dataset, ref_rows, qry_rows) -> Dataset:
    """Copy of `ds` whose listed reference / query rows store no values (an
    all-zero cell: its ranks are constant, so its correlations are NaN)."""
    def strip(data, indices, indptr, rows):
        rows = set(int(r) for r in rows)
        keep = np.ones(data.shape[0], dtype=bool)
        for r in rows:
            keep[indptr[r]:indptr[r + 1]] = False
        counts = np.diff(indptr)
        for r in rows:
            counts[r] = 0
        new_ptr = np.zeros_like(indptr)
        new_ptr[1:] = np.cumsum(counts)
        return data[keep], indices[keep], new_ptr

    rd, ri, rp = strip(ds.ref_data, ds.ref_indices, ds.ref_indptr, ref_rows)
    qd, qi, qp = strip(ds.qry_data, ds.qry_indices, ds.qry_indptr, qry_rows)
    return dataclasses.replace(ds, ref_data=rd, ref_indices=ri, ref_indptr=rp, qry_data=qd,
                               qry_indices=qi, qry_indptr=qp)
