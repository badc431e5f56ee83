"""The seeded synthetic generator: determinism and the configured shape."""

import numpy as np

from cusingler_amd import synth


def test_generator_is_deterministic():
    a = synth.generate("small", n_qry=200)
    b = synth.generate("small", n_qry=200)
    for k in ["ref_data", "ref_indices", "ref_indptr", "qry_data", "qry_indices", "qry_indptr",
              "celltype"]:
        assert np.array_equal(getattr(a, k), getattr(b, k))


def test_small_config_shape():
    ds = synth.generate("small")
    assert (ds.ref_height, ds.ref_width, ds.qry_height) == (1000, 2000, 2000)
    assert np.bincount(ds.celltype).tolist() == [100] * 10
    frac = ds.ref_data.shape[0] / (ds.ref_height * ds.ref_width)
    assert 0.06 < frac < 0.10
    assert np.all(ds.ref_data >= 1.0)  # log2(1 + count) with count >= 1
    assert np.all(np.diff(ds.ref_indptr) >= 1)


def test_pancreas_gene_layout():
    ds = synth.generate("pancreas", n_qry=100)
    shared = set(ds.ref_genes) & set(ds.qry_genes)
    assert len(shared) == 12000
    # shared genes keep their relative order on both sides
    r = [g for g in ds.ref_genes if g in shared]
    q = [g for g in ds.qry_genes if g in shared]
    assert r == q
