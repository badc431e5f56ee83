"""The product's host preprocessing (include/cusingler_prep.h) against the
reference's own src/io.cpp outputs held in the golden fixtures, against the
oracle, and its error behaviour."""

import glob
import os

import numpy as np
import pytest

from cusingler_amd import engine, synth
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _ds_from_npz(g):
    return synth.Dataset(cfg=None, ref_height=int(g["ref_height"]), ref_width=int(g["ref_width"]),
                         ref_data=g["ref_data"], ref_indices=g["ref_indices"],
                         ref_indptr=g["ref_indptr"], qry_height=int(g["qry_height"]),
                         qry_width=int(g["qry_width"]), qry_data=g["qry_data"],
                         qry_indices=g["qry_indices"], qry_indptr=g["qry_indptr"],
                         codes=list(g["codes"]), celltype=g["celltype"],
                         cellnames=list(g["cellnames"]), ref_genes=list(g["ref_genes"]),
                         qry_genes=list(g["qry_genes"]), qry_truth=np.zeros(0, np.int32))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))))
def test_prep_matches_reference_io(path):
    g = np.load(path)
    p = engine.Prep(_ds_from_npz(g), 4)
    assert np.array_equal(p.table(2), g["ref_common_genes"])
    assert np.array_equal(p.table(3), g["ref_thre_genes"])
    assert np.array_equal(p.table(0), g["ref_train_idxs"])
    assert np.array_equal(p.table(1), g["ref_train_values"])
    for step in (0, 1):
        c = p.generate(step)
        for k in ["ref_data", "ref_indices", "ref_indptr", "qry_data", "qry_indices", "qry_indptr"]:
            assert np.array_equal(getattr(c, k), g[f"step{step}_{k}"]), (step, k)
        assert c.max_uniq_gene == int(g["ref_max_uniq%d" % step])
    c1 = p.generate(1)
    assert np.array_equal(c1.ctdidx, g["ref_ctdidx1"])
    assert np.array_equal(c1.ctdiff, g["ref_ctdiff1"])


@pytest.mark.parametrize("name", ["small", "pancreas"])
def test_prep_matches_oracle(name):
    ds = synth.generate(name, n_qry=500)
    p = engine.Prep(ds, 8)
    o = O.Oracle(ds, 8)
    t = o.tables()
    for step in (0, 1):
        c = p.generate(step)
        oc = o.coded(step)
        for k, v in oc.items():
            assert np.array_equal(getattr(c, k), v), (step, k)
    assert np.array_equal(p.generate(1).ctdiff, t["ctdiff1"])


@pytest.mark.skipif(not O.ref_driver_available(), reason="oracle/_ref not built here")
def test_prep_matches_live_reference_build(tmp_path):
    ds = synth.generate("small", n_qry=300)
    ref = O.run_ref_driver(ds, str(tmp_path))
    p = engine.Prep(ds, 8)
    for step in (0, 1):
        c = p.generate(step)
        for k in ["ref_data", "ref_indices", "ref_indptr", "qry_data", "qry_indices", "qry_indptr"]:
            assert np.array_equal(getattr(c, k), ref[f"step{step}_{k}"])


def test_prep_error_cases():
    ds = synth.tiny()
    bad = synth.tiny()
    bad.qry_genes = [g + "_x" for g in bad.qry_genes]  # no shared gene
    with pytest.raises(engine.CsgError, match="No intersection"):
        engine.Prep(bad, 2)
    empty_label = synth.tiny()
    empty_label.celltype = np.where(empty_label.celltype == 2, 1, empty_label.celltype).astype(np.int32)
    with pytest.raises(engine.CsgError, match="no reference cell"):
        engine.Prep(empty_label, 2)
    out_of_range = synth.tiny()
    out_of_range.celltype = out_of_range.celltype.copy()
    out_of_range.celltype[0] = 7
    with pytest.raises(engine.CsgError, match="no name"):
        engine.Prep(out_of_range, 2)
    nan = synth.tiny()
    nan.ref_data = nan.ref_data.copy()
    nan.ref_data[3] = np.nan
    with pytest.raises(engine.CsgError, match="NaN"):
        engine.Prep(nan, 2)
    few_genes = synth.tiny(n_genes=100)  # fewer genes than the 263 per pair for 3 labels
    with pytest.raises(engine.CsgError, match="fewer reference genes"):
        engine.Prep(few_genes, 2)
    assert engine.Prep(ds, 2).filtered is False
