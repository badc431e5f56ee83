"""Reference training on the device (csg_train, src/io.cpp:141-262): the
pair windows, thre_genes and common_genes bit-identical to the reference's own
io.cpp outputs (golden fixtures, stored zeros and signed values among them)
and to the host preprocessing on the configurations' references (filtered
genes, the atlas, the 100-label streaming reference, dense rows).

Every case prints its name before touching the GPU."""

import glob
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from cusingler_amd import engine
    e = engine.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))))
def test_train_matches_reference_io(eng, path):
    from cusingler_amd import engine
    from cusingler_amd.synth import Dataset
    g = np.load(path)
    ds = Dataset(cfg=None, ref_height=int(g["ref_height"]), ref_width=int(g["ref_width"]),
                 ref_data=g["ref_data"], ref_indices=g["ref_indices"], ref_indptr=g["ref_indptr"],
                 qry_height=int(g["qry_height"]), qry_width=int(g["qry_width"]),
                 qry_data=g["qry_data"], qry_indices=g["qry_indices"], qry_indptr=g["qry_indptr"],
                 codes=list(g["codes"]), celltype=g["celltype"], cellnames=list(g["cellnames"]),
                 ref_genes=list(g["ref_genes"]), qry_genes=list(g["qry_genes"]),
                 qry_truth=np.zeros(0, np.int32))
    prep = engine.Prep(ds, 2)
    _say("train vs io.cpp", os.path.basename(path))
    t = eng.train(ds, prep.ref_map(ds.ref_width))
    assert np.array_equal(t["train_idxs"], g["ref_train_idxs"])
    assert np.array_equal(t["train_values"], g["ref_train_values"])
    assert np.array_equal(t["common_genes"], g["ref_common_genes"])
    assert np.array_equal(t["thre_genes"], g["ref_thre_genes"])


def _data(name):
    import dataclasses
    from cusingler_amd import synth
    if name == "wide4k":
        return synth.generate(synth.SynthConfig("wide4k", 40 * 5, 60, 20000, 0, 0, 40, 20, 2.0,
                                                6.0, 0.4, seed=3, label_sigma=2.0,
                                                marker_pool=1.0))
    cfg = dataclasses.replace(synth.CONFIGS[name], n_qry=64, n_train=0)
    return synth.generate(cfg)


@pytest.mark.parametrize("name", ["small", "pancreas", "atlas", "streaming", "wide4k"])
def test_train_matches_host(eng, name):
    from cusingler_amd import engine
    ds = _data(name)
    t0 = time.perf_counter()
    prep = engine.Prep(ds, 16)
    host_s = time.perf_counter() - t0
    _say("train vs host", name, ds.ref_height, "x", ds.ref_width, "labels", len(ds.codes))
    eng.train(ds, prep.ref_map(ds.ref_width))  # warm (code objects, allocator)
    t0 = time.perf_counter()
    t = eng.train(ds, prep.ref_map(ds.ref_width))
    dev_s = time.perf_counter() - t0
    _say(f"  host preprocessing (training + query setup, 16 threads) {host_s:.3f} s, "
         f"device training {dev_s:.3f} s")
    for k, name_ in enumerate(["train_idxs", "train_values", "common_genes", "thre_genes"]):
        assert np.array_equal(t[name_], prep.table(k)), name_
