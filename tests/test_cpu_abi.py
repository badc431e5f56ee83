"""The C-ABI library loads and exports every entry point the headers declare;
the drop-in Python module keeps the reference's signature; without a GPU the
product fails loudly instead of falling back."""

import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ["cusingler_amd.h", "cusingler_prep.h"]:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(csg_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(engine_lib):
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(engine_lib, n)]
    assert not missing, missing


def test_pymodule_signature_matches_reference():
    from cusingler import cusingler
    doc = cusingler.run.__doc__
    for arg in ["cores", "gpuid", "ref_height", "ref_width", "ref_data", "ref_indices",
                "ref_indptr", "qry_height", "qry_width", "qry_data", "qry_indices", "qry_indptr",
                "codes", "celltype", "cellnames", "ref_geneidx", "qry_geneidx"]:
        assert arg in doc
    import re as _re
    assert _re.search(r"quantile: [^=,]* = 80\b", doc)
    assert _re.search(r"finetune_thre: [^=,]* = 0.05\b", doc)
    assert _re.search(r"finetune_times: [^=,]* = 0\b", doc)
    # positional order is the reference's (src/main.cpp:84-88)
    order = ["cores", "gpuid", "ref_height", "ref_width", "ref_data", "ref_indices", "ref_indptr",
             "qry_height", "qry_width", "qry_data", "qry_indices", "qry_indptr", "codes",
             "celltype", "cellnames", "ref_geneidx", "qry_geneidx", "quantile", "finetune_thre",
             "finetune_times"]
    pos = [doc.index(a + ":") for a in order]
    assert pos == sorted(pos)


def _gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_gpu(), reason="checks the no-GPU behaviour")
def test_no_gpu_fails_loudly():
    from cusingler import cusingler
    from cusingler_amd import engine, synth
    with pytest.raises(engine.CsgError):
        engine.Engine(0)
    ds = synth.tiny()
    with pytest.raises(RuntimeError, match="gpu"):
        cusingler.run(2, 0, ds.ref_height, ds.ref_width, ds.ref_data.tolist(),
                      ds.ref_indices.tolist(), ds.ref_indptr.astype(int).tolist(), ds.qry_height,
                      ds.qry_width, ds.qry_data.tolist(), ds.qry_indices.tolist(),
                      ds.qry_indptr.astype(int).tolist(), ds.codes, ds.celltype.tolist(),
                      ds.cellnames, ds.ref_genes, ds.qry_genes, verbose=False)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
