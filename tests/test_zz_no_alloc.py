"""The allocation-free value coder (csg_code_csr_noalloc) against the
allocating one (csg_code_csr): two calls in a row on one engine, the second
with a smaller input, each bit-equal in codes, columns, row pointers and
max_uniq_gene; a NaN input raises the status flag; more rows than the fixed
work memory holds is refused.  Device memory for inputs and outputs comes from
torch (the library allocates none in the call)."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _case(ds, prep, rows):
    import torch
    dev = torch.device("cuda:0")
    ptr = np.asarray(ds.qry_indptr[:rows + 1], np.int64)
    nnz = int(ptr[-1])
    t = {
        "data": torch.from_numpy(np.ascontiguousarray(ds.qry_data[:nnz], np.float32)).to(dev),
        "indices": torch.from_numpy(np.ascontiguousarray(ds.qry_indices[:nnz], np.int32)).to(dev),
        "indptr": torch.from_numpy(ptr.astype(np.int32)).to(dev),
        "map_a": torch.from_numpy(prep.query_map(0, ds.qry_width)).to(dev),
        "map_b": torch.from_numpy(prep.query_map(1, ds.qry_width)).to(dev),
        "status": torch.full((4,), 7, dtype=torch.int32, device=dev),
    }
    for tag in ("x", "y"):  # x: allocating coder, y: allocation-free coder
        for s in ("a", "b"):
            t[f"{tag}data_{s}"] = torch.full((max(nnz, 1),), -1, dtype=torch.int16, device=dev)
            t[f"{tag}indices_{s}"] = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device=dev)
            t[f"{tag}indptr_{s}"] = torch.full((rows + 1,), -1, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    return t


def _out(t, tag):
    return [(t[f"{tag}data_{s}"].data_ptr(), t[f"{tag}indices_{s}"].data_ptr(),
             t[f"{tag}indptr_{s}"].data_ptr()) for s in ("a", "b")]


def _coded(t, tag, s):
    p = t[f"{tag}indptr_{s}"].cpu().numpy()
    n = int(p[-1])
    return (t[f"{tag}data_{s}"][:n].cpu().numpy().view(np.uint16),
            t[f"{tag}indices_{s}"][:n].cpu().numpy(), p)


def test_noalloc_coder_matches_allocating_coder():
    from cusingler_amd import engine, synth
    ds = synth.generate("pancreas", n_qry=1200)
    prep = engine.Prep(ds, 8)
    c0, c1 = prep.generate(0), prep.generate(1)
    eng = engine.Engine(0)
    try:
        for rows in (1200, 333):   # the second call on a smaller input
            t = _case(ds, prep, rows)
            _say("no-alloc coder, rows", rows)
            xa, xb = _out(t, "x")
            mu = eng.code_csr(rows, ds.qry_width, t["data"].data_ptr(), t["indices"].data_ptr(),
                              t["indptr"].data_ptr(), t["map_a"].data_ptr(), t["map_b"].data_ptr(),
                              xa, xb)
            ya, yb = _out(t, "y")
            eng.code_csr_noalloc(rows, ds.qry_width, t["data"].data_ptr(), t["indices"].data_ptr(),
                                 t["indptr"].data_ptr(), t["map_a"].data_ptr(),
                                 t["map_b"].data_ptr(), ya, yb, t["status"].data_ptr())
            eng.synchronize()   # the call itself does not wait
            st = t["status"].cpu().numpy()
            assert st[0] == 0 and st[3] == 0, st
            assert (int(st[1]), int(st[2])) == mu, (st, mu)
            for s, c in (("a", c0), ("b", c1)):
                dx, ix, px = _coded(t, "x", s)
                dy, iy, py = _coded(t, "y", s)
                assert np.array_equal(px, py) and np.array_equal(ix, iy), (rows, s)
                assert np.array_equal(dx, dy), (rows, s)
                # and the host preprocessing's codes of those rows
                assert np.array_equal(py, c.qry_indptr[:rows + 1]), (rows, s)
                assert np.array_equal(dy, c.qry_data[:py[-1]]), (rows, s)
        # a NaN value raises the status flag (the call itself returns)
        t = _case(ds, prep, 50)
        ind = ds.qry_indices[:int(ds.qry_indptr[50])]
        kept = int(np.nonzero(prep.query_map(0, ds.qry_width)[ind] >= 0)[0][0])
        t["data"][kept] = float("nan")  # a kept value (unkept ones are ignored)
        import torch
        torch.cuda.synchronize()  # the write is on torch's stream
        ya, yb = _out(t, "y")
        _say("no-alloc coder, NaN flag")
        eng.code_csr_noalloc(50, ds.qry_width, t["data"].data_ptr(), t["indices"].data_ptr(),
                             t["indptr"].data_ptr(), t["map_a"].data_ptr(), t["map_b"].data_ptr(),
                             ya, yb, t["status"].data_ptr())
        eng.synchronize()
        assert t["status"].cpu().numpy()[0] & 1
        # beyond the fixed work memory: refused before any launch
        with pytest.raises(engine.CsgError) as e:
            eng.code_csr_noalloc((1 << 20) + 1, ds.qry_width, t["data"].data_ptr(),
                                 t["indices"].data_ptr(), t["indptr"].data_ptr(),
                                 t["map_a"].data_ptr(), t["map_b"].data_ptr(), ya, yb,
                                 t["status"].data_ptr())
        assert e.value.status == 1
    finally:
        eng.close()
