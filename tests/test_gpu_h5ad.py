"""The .h5ad harness on the GPU against the oracle.

- ``h5ad.main`` (the reference's test/test_basic.py flow: load both files,
  ``cusingler.run``, TSV) on a pancreas-shaped dataset whose reference and
  query hold genes the other lacks, stored as the harness reads it: its first
  and final label names are the oracle's.
- ``h5ad.annotate_h5ad`` streams the same query file in ragged batches
  (hyperslab reads into page-locked buffers, device value coding, upload of the
  next batch during the current one) on the plain and exact paths, from
  uncompressed float32 and deflated float64 files: the oracle's labels, and the
  same TSV as ``main``.

Every case prints its name before touching the GPU."""

import dataclasses
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    from cusingler_amd import h5ad, synth
    from oracle import oracle as O
    try:
        h5ad.hdf5()
    except h5ad.H5Error as e:
        pytest.skip(str(e))
    cfg = dataclasses.replace(synth.CONFIGS["pancreas"], n_qry=1500)
    ds = synth.generate(cfg)
    d = tmp_path_factory.mktemp("h5ad")
    files = {"f32": (str(d / "ref.h5ad"), str(d / "qry.h5ad")),
             "f64_gzip": (str(d / "ref64.h5ad"), str(d / "qry64.h5ad"))}
    h5ad.write_dataset(ds, *files["f32"])
    ds64 = dataclasses.replace(ds, ref_data=ds.ref_data.astype(np.float64),
                               qry_data=ds.qry_data.astype(np.float64),
                               ref_indices=ds.ref_indices.astype(np.int64),
                               qry_indices=ds.qry_indices.astype(np.int64))
    h5ad.write_dataset(ds64, *files["f64_gzip"], compression=4, chunk=20000)
    orc = O.Oracle(ds, 16)
    of, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    ofin, _ = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    names = [f"cell_{i}" for i in range(ds.qry_height)] if len(ds.cellnames) != ds.qry_height \
        else list(ds.cellnames)
    want = [[c, ds.codes[a], ds.codes[b]] for c, a, b in zip(names, of, ofin)]
    return ds, files, want, d


def _read_tsv(path):
    with open(path) as fh:
        lines = fh.read().split("\n")
    assert lines[0] == "cell\tfirstLabel\tfinalLabel" and lines[-1] == ""
    return [ln.split("\t") for ln in lines[1:-1]]


def test_harness_main_matches_oracle(data):
    from cusingler_amd import h5ad
    ds, files, want, d = data
    _say("h5ad.main: pancreas-shaped,", ds.qry_height, "query cells, genes",
         ds.ref_width, "/", ds.qry_width)
    out = str(d / "main.tsv")
    res = h5ad.main(*files["f32"], out, cores=16, verbose=False)
    assert [list(r) for r in res] == want
    assert _read_tsv(out) == want


@pytest.mark.parametrize("layout,path,batch", [("f32", "plain", 512), ("f64_gzip", "plain", 700),
                                               ("f32", "exact_int8", 1500)])
def test_streaming_annotate_matches_oracle(data, layout, path, batch):
    from cusingler_amd import h5ad
    ds, files, want, d = data
    _say("annotate_h5ad:", layout, path, "batches of", batch)
    out = str(d / f"stream_{layout}_{path}.tsv")
    r = h5ad.annotate_h5ad(*files[layout], out, path=path, batch_rows=batch, cores=16)
    assert r["batches"] == (ds.qry_height + batch - 1) // batch
    assert _read_tsv(out) == want
