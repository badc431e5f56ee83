"""The reference's pipeline order run through the compiled shim of its device
API (cusingler_amd/shim): initGPU, copyin_score, get_label, destroy_score,
copyin, cufinetune, destroy -- first and final labels against the oracle, on
the plain and the exact INT8 paths.

Every case prints its name before touching the GPU."""

import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cusingler_amd", "shim", "shim_driver")


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _write(ds, d, cores=4):
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{ds.ref_height} {ds.ref_width} {ds.qry_height} {ds.qry_width} "
                f"{len(ds.codes)} {cores}\n")
    for name, a, dt in [("ref_data.f32", ds.ref_data, np.float32),
                        ("ref_indices.i32", ds.ref_indices, np.int32),
                        ("ref_indptr.i32", ds.ref_indptr, np.int32),
                        ("qry_data.f32", ds.qry_data, np.float32),
                        ("qry_indices.i32", ds.qry_indices, np.int32),
                        ("qry_indptr.i32", ds.qry_indptr, np.int32),
                        ("celltype.i32", ds.celltype, np.int32)]:
        np.ascontiguousarray(a, dt).tofile(os.path.join(d, name))
    for name, lst in [("ref_genes.txt", ds.ref_genes), ("qry_genes.txt", ds.qry_genes)]:
        with open(os.path.join(d, name), "w") as f:
            f.write("\n".join(lst) + "\n")


@pytest.mark.parametrize("name,path", [("tiny_filtered", 0), ("small", 0), ("small", 1),
                                       ("pancreas", 1)])
def test_shim_pipeline_matches_oracle(name, path):
    from cusingler_amd import synth
    from oracle import oracle as O
    ds = {"tiny_filtered": lambda: synth.tiny(n_ref_only=7, n_qry_only=5, seed=11),
          "small": lambda: synth.generate("small", n_qry=800),
          "pancreas": lambda: synth.generate("pancreas", n_qry=800)}[name]()
    orc = O.Oracle(ds, 8)
    of, om, _ = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
    ofin, _ = orc.finetune(0, ds.qry_height, om, 0.8, 0.05, 0, O.VAR_SEQ)
    with tempfile.TemporaryDirectory() as d:
        _write(ds, d)
        _say("shim pipeline", name, "path", path)
        r = subprocess.run([DRIVER, d, d, str(path)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        first = np.fromfile(os.path.join(d, "first.i32"), np.int32)
        final = np.fromfile(os.path.join(d, "final.u32"), np.uint32)
    assert np.array_equal(first, of), int((first != of).sum())
    assert np.array_equal(final, ofin), int((final != ofin).sum())
