"""The query column map of the device value coder (csg_prep_query_map), and a
NumPy restatement of csrFilter (src/io.cpp:470-523) driven by it, against the
reference's own io.cpp outputs held in the golden fixtures: the coded query
CSR of both steps and max_uniq_gene.  A planted mistake (ranks among all
stored values instead of the kept ones) must fail the same check."""

import glob
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dataset(g):
    from cusingler_amd.synth import Dataset
    return Dataset(cfg=None, ref_height=int(g["ref_height"]), ref_width=int(g["ref_width"]),
                   ref_data=g["ref_data"], ref_indices=g["ref_indices"], ref_indptr=g["ref_indptr"],
                   qry_height=int(g["qry_height"]), qry_width=int(g["qry_width"]),
                   qry_data=g["qry_data"], qry_indices=g["qry_indices"], qry_indptr=g["qry_indptr"],
                   codes=list(g["codes"]), celltype=g["celltype"], cellnames=list(g["cellnames"]),
                   ref_genes=list(g["ref_genes"]), qry_genes=list(g["qry_genes"]),
                   qry_truth=np.zeros(0, np.int32))


def csr_filter(data, indices, indptr, colmap, rank_over_kept=True):
    """csrFilter: kept entries in row order, columns through colmap, values
    replaced by their 1-based rank among the row's distinct kept values;
    returns (codes, columns, row pointers, max over rows of #distinct + 1)."""
    out_d, out_i, out_p = [], [], [0]
    mu = 0
    for r in range(len(indptr) - 1):
        a, b = indptr[r], indptr[r + 1]
        m = colmap[indices[a:b]]
        keep = m >= 0
        vals = data[a:b][keep]
        pool = vals if rank_over_kept else data[a:b]
        uniq = np.unique(pool)
        mu = max(mu, len(np.unique(vals)) + 1)
        out_d.append((np.searchsorted(uniq, vals) + 1).astype(np.uint16))
        out_i.append(m[keep].astype(np.int32))
        out_p.append(out_p[-1] + int(keep.sum()))
    cat = (lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt))
    return cat(out_d, np.uint16), cat(out_i, np.int32), np.array(out_p, np.int32), mu


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))))
def test_query_map_and_coding_match_reference_io(path):
    from cusingler_amd import engine
    g = np.load(path)
    ds = _dataset(g)
    prep = engine.Prep(ds, 2)
    for step in (0, 1):
        cmap = prep.query_map(step, ds.qry_width)
        d, i, p, mu = csr_filter(ds.qry_data, ds.qry_indices, ds.qry_indptr.astype(np.int64), cmap)
        assert np.array_equal(d, g[f"step{step}_qry_data"]), (path, step)
        assert np.array_equal(i, g[f"step{step}_qry_indices"]), (path, step)
        assert np.array_equal(p, g[f"step{step}_qry_indptr"]), (path, step)
        # max_uniq_gene of the step is the larger of the reference's and the query's
        rc = g[f"step{step}_ref_data"]
        rp = g[f"step{step}_ref_indptr"]
        ref_mu = max([len(np.unique(rc[rp[r]:rp[r + 1]])) + 1 for r in range(len(rp) - 1)])
        assert max(ref_mu, mu) == int(g[f"ref_max_uniq{step}"]), (path, step)


def test_planted_mistake_is_caught():
    """Ranking among every stored value of the row (not only the kept ones)
    is a plausible slip; it changes the codes of the filtered fixture."""
    from cusingler_amd import engine
    g = np.load(os.path.join(GOLDEN, "tiny_filtered.npz"))
    ds = _dataset(g)
    cmap = engine.Prep(ds, 2).query_map(0, ds.qry_width)
    d, _, _, _ = csr_filter(ds.qry_data, ds.qry_indices, ds.qry_indptr.astype(np.int64), cmap,
                            rank_over_kept=False)
    assert not np.array_equal(d, g["step0_qry_data"])


def test_query_map_shared_genes_is_the_gene_set_index():
    """Without filtering, the map of step s sends gene-set member k to k and
    every other column to -1."""
    from cusingler_amd import engine
    g = np.load(os.path.join(GOLDEN, "tiny_shared.npz"))
    ds = _dataset(g)
    prep = engine.Prep(ds, 2)
    for step, key in ((0, "ref_common_genes"), (1, "ref_thre_genes")):
        cmap = prep.query_map(step, ds.qry_width)
        genes = g[key]
        want = np.full(ds.qry_width, -1, np.int32)
        want[genes] = np.arange(len(genes))
        assert np.array_equal(cmap, want)
