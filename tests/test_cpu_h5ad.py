"""The .h5ad reader and writer (cusingler_amd/h5ad.py) on the CPU.

Pinned from outside this module where the image allows it:
- files written by HDF5's own ``h5import`` tool (big-endian float64 values
  with deflate, int64 indices, int32 row pointers, variable-length ASCII
  strings) read back as the arrays they were made from;
- files this module writes, read by HDF5's own ``h5dump`` tool: the AnnData
  attributes, the types of every dataset and their values;
and otherwise by round trips through the layouts the reference's harness
(test/test_basic.py:7-58) accepts: /obs/celltype, /obs/ClusterName,
/obsm/annotation_au/celltype, /var's index column and its Gene_ID -> Symbol
rule, plus the TSV table against numpy.savetxt (test/test_basic.py:71-73)."""

import os
import re
import subprocess

import numpy as np
import pytest

h5ad = pytest.importorskip("cusingler_amd.h5ad")


def _hdf5_or_skip():
    try:
        h5ad.hdf5()
    except h5ad.H5Error as e:
        pytest.skip(str(e))


def _tool(name):
    for d in (os.environ.get("PATH", "").split(os.pathsep) + ["/opt/conda/bin"]):
        p = os.path.join(d, name)
        if d and os.access(p, os.X_OK):
            return p
    pytest.skip(f"{name} not found")


def _csr(rng, rows, cols, density, empty_rows=()):
    m = (rng.random((rows, cols)) < density) * rng.integers(1, 50, (rows, cols))
    for r in empty_rows:
        m[r] = 0
    vals = np.log2(1 + m.astype(np.float64))
    indptr = np.zeros(rows + 1, np.int64)
    data, ind = [], []
    for r in range(rows):
        nz = np.nonzero(m[r])[0]
        ind.append(nz)
        data.append(vals[r, nz])
        indptr[r + 1] = indptr[r] + nz.size
    return np.concatenate(data), np.concatenate(ind).astype(np.int64), indptr


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(7)


def _import(tmp, out, specs):
    """h5import: one (path, array or list of str, options) per dataset."""
    tool = _tool("h5import")
    args = [tool]
    for k, (path, a, opts) in enumerate(specs):
        cfg = os.path.join(tmp, f"in{k}.cfg")
        if isinstance(a, list):
            src = os.path.join(tmp, f"in{k}.txt")
            with open(src, "w") as fh:
                fh.write("\n".join(a) + "\n")
            lines = [f"PATH {path}", "INPUT-CLASS STR"]
        else:
            src = os.path.join(tmp, f"in{k}.bin")
            a.tofile(src)  # native order, as h5import reads binary input
            cls = "FP" if a.dtype.kind == "f" else "IN"
            lines = [f"PATH {path}", f"INPUT-CLASS {cls}", f"INPUT-SIZE {a.dtype.itemsize * 8}",
                     "RANK 1", f"DIMENSION-SIZES {a.size}", f"OUTPUT-CLASS {cls}"] + opts
        with open(cfg, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        args += [src, "-c", cfg]
    args += ["-o", out]
    subprocess.run(args, check=True, capture_output=True)


def test_reads_files_written_by_h5import(tmp_path, rng):
    """An h5ad whose datasets h5import wrote (only the attributes the harness
    reads are added by this module): every loader returns the source arrays,
    converted as the reference's binding converts them."""
    _hdf5_or_skip()
    data, ind, ptr = _csr(rng, 40, 30, 0.3, empty_rows=(0, 17, 39))
    names = [f"c{i}_x" for i in range(40)]
    genes = [f"G{i:03d}" for i in range(30)]
    cats = ["B cell", "T cell", "NK"]
    codes = rng.integers(0, 3, 40).astype(np.int8)
    f = str(tmp_path / "imp.h5ad")
    _import(str(tmp_path), f, [
        ("X/data", data, ["OUTPUT-SIZE 64", "OUTPUT-ARCHITECTURE IEEE", "OUTPUT-BYTE-ORDER BE",
                          "CHUNKED-DIMENSION-SIZES 16", "COMPRESSION-TYPE GZIP",
                          "COMPRESSION-PARAM 6"]),
        ("X/indices", ind, ["OUTPUT-SIZE 64", "OUTPUT-ARCHITECTURE STD", "OUTPUT-BYTE-ORDER LE",
                            "CHUNKED-DIMENSION-SIZES 7"]),
        ("X/indptr", ptr.astype(np.int32), ["OUTPUT-SIZE 32", "OUTPUT-ARCHITECTURE STD",
                                            "OUTPUT-BYTE-ORDER BE"]),
        ("obs/_index", names, []),
        ("obs/celltype/categories", cats, []),
        ("obs/celltype/codes", codes, ["OUTPUT-SIZE 8"]),
        ("var/_index", genes, []),
    ])
    with h5ad.H5File(f, "a") as fh:
        fh.set_attr("/X", "shape", np.array([40, 30], np.int64))
        fh.set_attr("/var", "_index", "_index")
        assert fh.dtype("/X/data") == np.float64 and fh.dtype("/X/indices") == np.int64
        assert fh.dtype("/X/indptr") == np.int32 and fh.dtype("/obs/celltype/codes") == np.int8
        assert np.array_equal(fh.read("/X/data", 5, 23), data[5:23])
    h, w, d, i, p = h5ad.load_matrix(f)
    assert (h, w) == (40, 30)
    assert d.dtype == np.float32 and np.array_equal(d, data.astype(np.float32))
    assert i.dtype == np.int32 and np.array_equal(i, ind)
    assert np.array_equal(p, ptr)
    got_names, got_codes = h5ad.load_celltypes(f)
    assert got_names == cats and np.array_equal(got_codes, codes)
    assert h5ad.load_cellnames(f) == names
    assert h5ad.load_geneidx(f) == genes
    rows = h5ad.CsrRows(f)
    bd, bi, bp = np.empty(ptr[-1], np.float32), np.empty(ptr[-1], np.int32), np.empty(41, np.int32)
    for r0, r1 in ((0, 1), (0, 40), (3, 18), (17, 18), (39, 40)):
        n = rows.read(r0, r1, bd, bi, bp)
        assert n == ptr[r1] - ptr[r0]
        assert np.array_equal(bd[:n], data[ptr[r0]:ptr[r1]].astype(np.float32))
        assert np.array_equal(bi[:n], ind[ptr[r0]:ptr[r1]])
        assert np.array_equal(bp[:r1 - r0 + 1], ptr[r0:r1 + 1] - ptr[r0])
    rows.close()


def _dump(tool, f, *args):
    return subprocess.run([tool, *args, f], check=True, capture_output=True, text=True).stdout


def test_written_layout_read_by_h5dump(tmp_path, rng):
    """What write_h5ad writes, as HDF5's h5dump reads it: AnnData's
    encoding attributes, X's shape, the dataset types, UTF-8 variable-length
    names and the values themselves."""
    _hdf5_or_skip()
    tool = _tool("h5dump")
    data, ind, ptr = _csr(rng, 12, 9, 0.4, empty_rows=(4,))
    f = str(tmp_path / "w.h5ad")
    h5ad.write_h5ad(f, 12, 9, data.astype(np.float32), ind.astype(np.int32), ptr,
                    [f"cell{i}" for i in range(12)], [f"g{i}" for i in range(9)],
                    obs={"celltype": (["a", "b"], np.arange(12) % 2)})
    attrs = _dump(tool, f, "-A")
    for want in ('"anndata"', '"csr_matrix"', '"dataframe"', '"categorical"', '"string-array"'):
        assert want in attrs, want
    shape = re.search(r'ATTRIBUTE "shape" \{\s*DATATYPE\s+H5T_STD_I64LE.*?DATA \{\s*\(0\): ([^}]*)\}',
                      attrs, re.S)
    assert shape and [int(x) for x in shape.group(1).split(",")] == [12, 9]
    types = {"/X/data": "H5T_IEEE_F32LE", "/X/indices": "H5T_STD_I32LE",
             "/X/indptr": "H5T_STD_I64LE", "/obs/celltype/codes": "H5T_STD_I8LE"}
    for path, t in types.items():
        # nine significant digits print every float32 exactly
        out = _dump(tool, f, "-d", path, "-y", "-w", "0", "-m", "%.9g")
        assert f"DATATYPE  {t}" in out, (path, out[:200])
        vals = re.search(r"DATA \{\s*(.*?)\s*\}", out, re.S).group(1)
        got = np.array([float(x) for x in vals.replace("\n", " ").split(",") if x.strip()])
        want = {"/X/data": data.astype(np.float32), "/X/indices": ind, "/X/indptr": ptr,
                "/obs/celltype/codes": np.arange(12) % 2}[path]
        assert np.array_equal(got.astype(np.float32), np.asarray(want).astype(np.float32)), path
    out = _dump(tool, f, "-d", "/obs/_index", "-y", "-w", "0")
    assert "STRSIZE H5T_VARIABLE" in out and "CSET H5T_CSET_UTF8" in out
    body = re.search(r"DATA \{\s*(.*?)\s*\}", out, re.S).group(1)
    assert re.findall(r'"([^"]*)"', body) == [f"cell{i}" for i in range(12)]


@pytest.mark.parametrize("variant", ["f32_i32", "f64_i64_gzip", "fixed_strings", "ragged_big"])
def test_round_trip(tmp_path, rng, variant):
    _hdf5_or_skip()
    rows, cols = (300, 50) if variant == "ragged_big" else (25, 40)
    data, ind, ptr = _csr(rng, rows, cols, 0.2, empty_rows=(0, rows - 1))
    kw = {}
    if variant == "f32_i32":
        data, ind = data.astype(np.float32), ind.astype(np.int32)
    elif variant == "f64_i64_gzip":
        kw = dict(compression=5, chunk=33)
    elif variant == "fixed_strings":
        kw = dict(fixed_strings=True)
    else:
        data = data.astype(np.float32)
        kw = dict(chunk=100, compression=1)
    names = [f"cellé{i}" for i in range(rows)]  # non-ASCII survives as UTF-8
    genes = [f"GENE-{i}" for i in range(cols)]
    cats = [f"type {k}" for k in range(7)]
    ct = rng.integers(0, 7, rows)
    f = str(tmp_path / f"{variant}.h5ad")
    h5ad.write_h5ad(f, rows, cols, data, ind, ptr, names, genes,
                    obs={"celltype": (cats, ct), "n_counts": np.arange(rows, dtype=np.float32)},
                    **kw)
    h, w, d, i, p = h5ad.load_matrix(f)
    assert (h, w) == (rows, cols)
    assert np.array_equal(d, np.asarray(data).astype(np.float32)) and np.array_equal(i, ind)
    assert np.array_equal(p, ptr)
    names_got, ct_got = h5ad.load_celltypes(f)
    assert names_got == cats and np.array_equal(ct_got, ct)
    assert h5ad.load_cellnames(f) == names
    assert h5ad.load_geneidx(f) == genes
    # row batches of every size, concatenated, give the whole matrix
    r = h5ad.CsrRows(f)
    for b in (1, 7, rows):
        bd, bi = np.empty(max(r.nnz, 1), np.float32), np.empty(max(r.nnz, 1), np.int32)
        bp = np.empty(b + 1, np.int32)
        cd, ci = [], []
        for r0 in range(0, rows, b):
            r1 = min(rows, r0 + b)
            n = r.read(r0, r1, bd, bi, bp)
            assert bp[0] == 0 and bp[r1 - r0] == n
            cd.append(bd[:n].copy())
            ci.append(bi[:n].copy())
        assert np.array_equal(np.concatenate(cd), d) and np.array_equal(np.concatenate(ci), i)
    r.close()


def test_harness_fallbacks(tmp_path, rng):
    """/obs/ClusterName when there is no /obs/celltype; the string dataset
    /obsm/annotation_au/celltype coded by first appearance; /var indexed by
    Gene_ID reads the Symbol column (test/test_basic.py:17-56)."""
    _hdf5_or_skip()
    data, ind, ptr = _csr(rng, 10, 6, 0.5)
    genes_id = [f"ENSG{i:05d}" for i in range(6)]
    symbols = [f"SYM{i}" for i in range(6)]
    f1 = str(tmp_path / "cluster.h5ad")
    h5ad.write_h5ad(f1, 10, 6, data, ind, ptr, [str(i) for i in range(10)], genes_id,
                    obs={"ClusterName": (["x", "y", "z"], np.arange(10) % 3)},
                    var={"Gene_ID": genes_id, "Symbol": symbols}, var_index_name="Gene_ID")
    names, codes = h5ad.load_celltypes(f1)
    assert names == ["x", "y", "z"] and np.array_equal(codes, np.arange(10) % 3)
    assert h5ad.load_geneidx(f1) == symbols
    f2 = str(tmp_path / "obsm.h5ad")
    ann = ["T", "B", "T", "NK", "B", "B", "T", "mono", "NK", "T"]
    h5ad.write_h5ad(f2, 10, 6, data, ind, ptr, [str(i) for i in range(10)], genes_id,
                    obsm_strings={"annotation_au/celltype": ann})
    names, codes = h5ad.load_celltypes(f2)
    assert names == ["T", "B", "NK", "mono"]
    assert codes.tolist() == [0, 1, 0, 2, 1, 1, 0, 3, 2, 0]
    assert h5ad.load_geneidx(f2) == genes_id
    f3 = str(tmp_path / "none.h5ad")
    h5ad.write_h5ad(f3, 10, 6, data, ind, ptr, [str(i) for i in range(10)], genes_id)
    with pytest.raises(h5ad.H5Error, match="no celltype"):
        h5ad.load_celltypes(f3)
    f4 = str(tmp_path / "missing_code.h5ad")
    h5ad.write_h5ad(f4, 10, 6, data, ind, ptr, [str(i) for i in range(10)], genes_id,
                    obs={"celltype": (["a", "b"], np.array([0, 1, -1, 0, 1, 0, 1, 0, 1, 0]))})
    with pytest.raises(h5ad.H5Error, match="code -1"):
        h5ad.load_celltypes(f4)


def test_errors(tmp_path, rng):
    _hdf5_or_skip()
    with pytest.raises(h5ad.H5Error, match="cannot open"):
        h5ad.load_matrix(str(tmp_path / "absent.h5ad"))
    bad = str(tmp_path / "bad.h5ad")
    with open(bad, "w") as fh:
        fh.write("not hdf5")
    with pytest.raises(h5ad.H5Error):
        h5ad.load_matrix(bad)
    data, ind, ptr = _csr(rng, 5, 4, 0.5)
    f = str(tmp_path / "dense.h5ad")
    with h5ad.H5File(f, "w") as fh:
        fh.write("/X", np.zeros((5, 4), np.float32))
    with pytest.raises(h5ad.H5Error, match="not a CSR group"):
        h5ad.load_matrix(f)


def test_tsv_matches_numpy_savetxt(tmp_path):
    """write_tsv's bytes equal the harness's np.savetxt call."""
    res = [["cell_0", "B cell", "B cell"], ["c1", "T", "NK"], ["x y", "a", "b"]]
    a, b = str(tmp_path / "a.tsv"), str(tmp_path / "b.tsv")
    h5ad.write_tsv(a, res)
    np.savetxt(b, np.array(res), delimiter="\t", header="cell\tfirstLabel\tfinalLabel",
               comments="", fmt="%s")
    assert open(a, "rb").read() == open(b, "rb").read()


def test_uint16_columns(tmp_path, rng):
    """Row batches read with 16-bit columns (the library converting while
    reading) equal the int32 ones; a column outside uint16 fails the read
    instead of being clamped."""
    _hdf5_or_skip()
    data, ind, ptr = _csr(rng, 30, 50, 0.3)
    ind = ind * 1000 + 7  # columns up to 49007, over int16
    f = str(tmp_path / "u16.h5ad")
    h5ad.write_h5ad(f, 30, 65536, data, ind.astype(np.int64), ptr, [str(i) for i in range(30)],
                    [f"g{i}" for i in range(65536)], compression=2, chunk=64)
    r = h5ad.CsrRows(f)
    n = int(ptr[-1])
    d16, i16, p16 = np.empty(n, np.float32), np.empty(n, np.uint16), np.empty(31, np.int32)
    d32, i32, p32 = np.empty(n, np.float32), np.empty(n, np.int32), np.empty(31, np.int32)
    assert r.read(0, 30, d16, i16, p16) == r.read(0, 30, d32, i32, p32) == n
    assert np.array_equal(i16.astype(np.int32), i32) and np.array_equal(d16, d32)
    assert np.array_equal(i32, ind) and np.array_equal(p16, p32)
    r.close()
    bad = ind.copy()
    bad[5] = 65536 + 5
    fb = str(tmp_path / "u16_bad.h5ad")
    h5ad.write_h5ad(fb, 30, 65536, data, bad.astype(np.int64), ptr, [str(i) for i in range(30)],
                    [f"g{i}" for i in range(65536)])
    r = h5ad.CsrRows(fb)
    with pytest.raises(h5ad.H5Error, match="outside"):
        r.read(0, 30, d16, i16, p16)
    r.close()
    fw = str(tmp_path / "wide.h5ad")
    h5ad.write_h5ad(fw, 30, 70000, data, ind, ptr, [str(i) for i in range(30)],
                    [f"g{i}" for i in range(70000)])
    r = h5ad.CsrRows(fw)
    with pytest.raises(ValueError, match="65536"):
        r.read(0, 30, d16, i16, p16)
    r.close()
