"""AnnData (``.h5ad``) input and TSV output around the annotation path.

The reference's harness (test/test_basic.py:7-58) reads the reference and query
matrices, the reference's cell types, the query's cell names and both gene
lists from ``.h5ad`` files with h5py, calls ``cusingler.run`` and writes
``cell / firstLabel / finalLabel`` as TSV (test/test_basic.py:71-73, the same
table as Pipeline::dump, src/pipeline.cpp:113-121).  This module does the
same over the HDF5 C library (``libhdf5``, 1.10 or later, through ctypes; h5py
is not needed):

- ``load_matrix``, ``load_celltypes``, ``load_cellnames``, ``load_geneidx``
  read what the harness's functions of those names read, with the same
  fallbacks (``/obs/celltype``, ``/obs/ClusterName``,
  ``/obsm/annotation_au/celltype``; ``/var``'s ``_index`` column, ``Symbol``
  when that is ``Gene_ID``);
- ``write_tsv`` writes the harness's table; ``main`` is the harness;
- ``CsrRows`` reads row batches of ``/X`` by hyperslab straight into caller
  buffers (page-locked ones for the GPU upload), and ``annotate_h5ad`` streams
  a query file through the device: each batch is read by a host thread while
  the previous batch uploads, is value-coded on the device and annotated;
- ``write_h5ad`` writes the AnnData on-disk layout (CSR ``/X``, ``obs`` /
  ``var`` dataframes with string indexes and categorical columns) for
  synthetic data and tests.

The harness swaps two names: it reads the categorical's ``codes`` (per-cell
integers) into ``codes`` and its ``categories`` (label names) into
``celltypes`` and passes them to ``run(..., codes, celltypes, ...)``, whose
parameters are the label names and the per-cell codes (src/main.cpp:75-76).
``load_celltypes`` returns ``(names, per-cell codes)``, the order ``run``
takes.
"""

from __future__ import annotations

import ctypes as C
import ctypes.util
import os
import sys
import threading
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# HDF5 C API constants (H5Fpublic.h, H5Ipublic.h, H5Tpublic.h, H5Spublic.h)
_ACC_RDONLY, _ACC_RDWR, _ACC_TRUNC = 0x0000, 0x0001, 0x0002
_P_DEFAULT = 0
_S_ALL = 0
_S_SCALAR = 0
_S_SELECT_SET = 0
_I_GROUP, _I_DATASET = 2, 5
_T_INTEGER, _T_FLOAT, _T_STRING, _T_ENUM = 0, 1, 3, 8
_T_SGN_NONE = 0
_T_CSET_UTF8 = 1
_T_VARIABLE = C.c_size_t(-1).value
_T_DIR_DEFAULT = 0

_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_ulonglong
_P = C.POINTER

_H5 = None


class H5Error(RuntimeError):
    pass


# H5T_conv_except_func_t returning H5T_CONV_ABORT (-1) for every exception
_CONV_CB = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p)
_ABORT_CB = _CONV_CB(lambda *args: -1)


def _candidates() -> List[str]:
    out = []
    env = os.environ.get("CSG_HDF5_LIB")
    if env:
        out.append(env)
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    out += ["libhdf5.so", "/opt/conda/lib/libhdf5.so",
            "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"]
    return out


def hdf5():
    """The HDF5 C library (first of $CSG_HDF5_LIB, the loader's search path,
    /opt/conda/lib), with the prototypes this module uses."""
    global _H5
    if _H5 is not None:
        return _H5
    lib, errs = None, []
    for cand in _candidates():
        try:
            lib = C.CDLL(cand)
            break
        except OSError as e:
            errs.append(f"{cand}: {e}")
    if lib is None:
        raise H5Error("the HDF5 C library (libhdf5) was not found; set CSG_HDF5_LIB to its "
                      "path (" + "; ".join(errs) + ")")
    lib.H5open.restype = _herr
    lib.H5open()
    maj, mi, rel = C.c_uint(), C.c_uint(), C.c_uint()
    lib.H5get_libversion(C.byref(maj), C.byref(mi), C.byref(rel))
    if (maj.value, mi.value) < (1, 10):
        raise H5Error(f"libhdf5 {maj.value}.{mi.value}.{rel.value} is older than 1.10 "
                      "(64-bit identifiers)")
    proto = {
        "H5Eset_auto2": (_herr, [_hid, C.c_void_p, C.c_void_p]),
        "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
        "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Lexists": (C.c_int, [_hid, C.c_char_p, _hid]),
        "H5Oopen": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Oclose": (_herr, [_hid]),
        "H5Iget_type": (C.c_int, [_hid]),
        "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
        "H5Gclose": (_herr, [_hid]),
        "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dclose": (_herr, [_hid]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dget_type": (_hid, [_hid]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dvlen_reclaim": (_herr, [_hid, _hid, _hid, C.c_void_p]),
        "H5Screate": (_hid, [C.c_int]),
        "H5Screate_simple": (_hid, [C.c_int, _P(_hsize), _P(_hsize)]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (C.c_int, [_hid, _P(_hsize), _P(_hsize)]),
        "H5Sselect_hyperslab": (_herr, [_hid, C.c_int, _P(_hsize), _P(_hsize), _P(_hsize),
                                        _P(_hsize)]),
        "H5Sclose": (_herr, [_hid]),
        "H5Tget_class": (C.c_int, [_hid]),
        "H5Tget_size": (C.c_size_t, [_hid]),
        "H5Tget_sign": (C.c_int, [_hid]),
        "H5Tget_super": (_hid, [_hid]),
        "H5Tget_native_type": (_hid, [_hid, C.c_int]),
        "H5Tis_variable_str": (C.c_int, [_hid]),
        "H5Tcopy": (_hid, [_hid]),
        "H5Tset_size": (_herr, [_hid, C.c_size_t]),
        "H5Tset_cset": (_herr, [_hid, C.c_int]),
        "H5Tget_cset": (C.c_int, [_hid]),
        "H5Tclose": (_herr, [_hid]),
        "H5Aexists": (C.c_int, [_hid, C.c_char_p]),
        "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
        "H5Aget_space": (_hid, [_hid]),
        "H5Aget_type": (_hid, [_hid]),
        "H5Aread": (_herr, [_hid, _hid, C.c_void_p]),
        "H5Awrite": (_herr, [_hid, _hid, C.c_void_p]),
        "H5Aclose": (_herr, [_hid]),
        "H5Pcreate": (_hid, [_hid]),
        "H5Pset_chunk": (_herr, [_hid, C.c_int, _P(_hsize)]),
        "H5Pset_deflate": (_herr, [_hid, C.c_uint]),
        "H5Pset_type_conv_cb": (_herr, [_hid, C.c_void_p, C.c_void_p]),
        "H5Pclose": (_herr, [_hid]),
    }
    for name, (res, args) in proto.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    # errors are reported by this module's exceptions, not printed by the library
    lib.H5Eset_auto2(0, None, None)

    def g(name):
        return _hid.in_dll(lib, name).value
    lib.native = {
        np.dtype(np.float32): g("H5T_NATIVE_FLOAT_g"), np.dtype(np.float64): g("H5T_NATIVE_DOUBLE_g"),
        np.dtype(np.int8): g("H5T_NATIVE_SCHAR_g"), np.dtype(np.uint8): g("H5T_NATIVE_UCHAR_g"),
        np.dtype(np.int16): g("H5T_NATIVE_SHORT_g"), np.dtype(np.uint16): g("H5T_NATIVE_USHORT_g"),
        np.dtype(np.int32): g("H5T_NATIVE_INT_g"), np.dtype(np.uint32): g("H5T_NATIVE_UINT_g"),
        np.dtype(np.int64): g("H5T_NATIVE_LLONG_g"), np.dtype(np.uint64): g("H5T_NATIVE_ULLONG_g"),
    }
    lib.c_s1 = g("H5T_C_S1_g")
    lib.dcpl_class = g("H5P_CLS_DATASET_CREATE_ID_g")
    # a transfer property list whose type conversions fail on any value
    # outside the destination type (the library clamps them by default)
    lib.strict_dxpl = lib.H5Pcreate(g("H5P_CLS_DATASET_XFER_ID_g"))
    _ok(lib.H5Pset_type_conv_cb(lib.strict_dxpl, C.cast(_ABORT_CB, C.c_void_p), None),
        "H5Pset_type_conv_cb")
    _H5 = lib
    return lib


def _ok(rc, what):
    if rc < 0:
        raise H5Error(f"HDF5: {what} failed")
    return rc


def _dims(h, space) -> Tuple[int, ...]:
    n = h.H5Sget_simple_extent_ndims(space)
    _ok(n, "H5Sget_simple_extent_ndims")
    if n == 0:
        return ()
    d = (_hsize * n)()
    _ok(h.H5Sget_simple_extent_dims(space, d, None), "H5Sget_simple_extent_dims")
    return tuple(int(x) for x in d)


def _np_dtype(h, ftype) -> np.dtype:
    """The numpy type a numeric HDF5 type reads into without loss."""
    cls = h.H5Tget_class(ftype)
    size = int(h.H5Tget_size(ftype))
    if cls == _T_FLOAT:
        return np.dtype(np.float64 if size > 4 else np.float32)
    if cls in (_T_INTEGER, _T_ENUM):
        base = h.H5Tget_super(ftype) if cls == _T_ENUM else ftype
        unsigned = h.H5Tget_sign(base) == _T_SGN_NONE
        if cls == _T_ENUM:
            h.H5Tclose(base)
        return np.dtype(("u" if unsigned else "i") + str(size))
    raise H5Error(f"not a numeric HDF5 type (class {cls})")


class H5File:
    """One open HDF5 file: object lookup, whole or partial dataset reads,
    attributes; in mode "w" (a new file) and "a" (an existing one) also
    groups, datasets and attributes."""

    def __init__(self, path: str, mode: str = "r"):
        self.h = hdf5()
        self.path = path
        if mode == "r":
            self.fid = self.h.H5Fopen(os.fsencode(path), _ACC_RDONLY, _P_DEFAULT)
        elif mode == "a":
            self.fid = self.h.H5Fopen(os.fsencode(path), _ACC_RDWR, _P_DEFAULT)
        elif mode == "w":
            self.fid = self.h.H5Fcreate(os.fsencode(path), _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT)
        else:
            raise ValueError(f"mode {mode!r}")
        if self.fid < 0:
            raise H5Error(f"cannot open {path!r} as HDF5 (mode {mode})")

    def close(self):
        if getattr(self, "fid", -1) >= 0:
            self.h.H5Fclose(self.fid)
            self.fid = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    # ---- lookup -----------------------------------------------------------
    def kind(self, path: str) -> Optional[str]:
        """"group", "dataset", another object kind, or None if absent (every
        link on the path is checked, as h5py's ``in`` does)."""
        parts = [p for p in path.split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if self.h.H5Lexists(self.fid, cur.encode(), _P_DEFAULT) <= 0:
                return None
        if not parts:
            return "group"
        oid = self.h.H5Oopen(self.fid, cur.encode(), _P_DEFAULT)
        if oid < 0:
            return None
        t = self.h.H5Iget_type(oid)
        self.h.H5Oclose(oid)
        return {_I_GROUP: "group", _I_DATASET: "dataset"}.get(t, "other")

    def __contains__(self, path: str) -> bool:
        return self.kind(path) is not None

    def _open_obj(self, path: str) -> int:
        oid = self.h.H5Oopen(self.fid, path.encode(), _P_DEFAULT)
        if oid < 0:
            raise KeyError(f"{self.path}: no object {path!r}")
        return oid

    def _open_dataset(self, path: str) -> int:
        did = self.h.H5Dopen2(self.fid, path.encode(), _P_DEFAULT)
        if did < 0:
            raise KeyError(f"{self.path}: no dataset {path!r}")
        return did

    # ---- datasets -----------------------------------------------------------
    def shape(self, path: str) -> Tuple[int, ...]:
        did = self._open_dataset(path)
        try:
            sp = self.h.H5Dget_space(did)
            try:
                return _dims(self.h, sp)
            finally:
                self.h.H5Sclose(sp)
        finally:
            self.h.H5Dclose(did)

    def dtype(self, path: str) -> np.dtype:
        """The numpy type a numeric dataset reads into without loss."""
        did = self._open_dataset(path)
        try:
            t = self.h.H5Dget_type(did)
            try:
                return _np_dtype(self.h, t)
            finally:
                self.h.H5Tclose(t)
        finally:
            self.h.H5Dclose(did)

    def read(self, path: str, start: int = 0, stop: Optional[int] = None, dtype=None,
             out: Optional[np.ndarray] = None, strict: bool = False) -> np.ndarray:
        """Elements [start, stop) of a 1-D numeric dataset (the whole of an n-D
        one when start = 0 and stop is None), converted by the library to
        `dtype` (default: the file's type without loss) and written into `out`
        when given (a contiguous array of that dtype and length).  `strict`:
        a value outside `dtype`'s range fails the read (the library would
        clamp it)."""
        h = self.h
        did = self._open_dataset(path)
        fsp = h.H5Dget_space(did)
        try:
            dims = _dims(h, fsp)
            ftype = h.H5Dget_type(did)
            try:
                dt = np.dtype(dtype) if dtype is not None else _np_dtype(h, ftype)
            finally:
                h.H5Tclose(ftype)
            if dt not in h.native:
                raise H5Error(f"{path}: no native HDF5 type for {dt}")
            if start == 0 and stop is None:
                shape = dims
                n = int(np.prod(dims)) if dims else 1
                msp, fsel = _S_ALL, _S_ALL
            else:
                if len(dims) != 1:
                    raise H5Error(f"{path}: partial reads need a 1-D dataset (shape {dims})")
                stop = dims[0] if stop is None else stop
                if not 0 <= start <= stop <= dims[0]:
                    raise IndexError(f"{path}: [{start}, {stop}) outside [0, {dims[0]})")
                n = stop - start
                shape = (n,)
                msp, fsel = None, fsp
            if out is None:
                out = np.empty(shape, dt)
            elif out.dtype != dt or out.size != n or not out.flags.c_contiguous:
                raise ValueError(f"{path}: `out` must be a contiguous {dt} array of {n} elements")
            if n == 0:
                return out
            xfer = h.strict_dxpl if strict else _P_DEFAULT
            if msp is None:
                st, cnt = (_hsize * 1)(start), (_hsize * 1)(n)
                _ok(h.H5Sselect_hyperslab(fsp, _S_SELECT_SET, st, None, cnt, None),
                    "H5Sselect_hyperslab")
                msp = h.H5Screate_simple(1, cnt, None)
                try:
                    _ok(h.H5Dread(did, h.native[dt], msp, fsel, xfer,
                                  out.ctypes.data_as(C.c_void_p)),
                        f"read {path}" + (f" (a value outside {dt})" if strict else ""))
                finally:
                    h.H5Sclose(msp)
            else:
                _ok(h.H5Dread(did, h.native[dt], _S_ALL, _S_ALL, xfer,
                              out.ctypes.data_as(C.c_void_p)),
                    f"read {path}" + (f" (a value outside {dt})" if strict else ""))
            return out
        finally:
            h.H5Sclose(fsp)
            h.H5Dclose(did)

    def read_strings(self, path: str) -> List[str]:
        """A 1-D (or scalar) dataset of fixed- or variable-length strings, as
        str (UTF-8, undecodable bytes replaced)."""
        did = self._open_dataset(path)
        try:
            return self._strings(did, self.h.H5Dget_space, self.h.H5Dget_type,
                                 lambda mt, buf: self.h.H5Dread(did, mt, _S_ALL, _S_ALL,
                                                                _P_DEFAULT, buf), path)
        finally:
            self.h.H5Dclose(did)

    def _strings(self, oid, get_space, get_type, read, what) -> List[str]:
        h = self.h
        sp, ft = get_space(oid), get_type(oid)
        try:
            dims = _dims(h, sp)
            n = int(np.prod(dims)) if dims else 1
            if h.H5Tget_class(ft) != _T_STRING:
                raise H5Error(f"{what}: not a string type")
            if n == 0:
                return []
            mt = h.H5Tcopy(h.c_s1)
            try:
                # the library converts no character set into another: read in
                # the file's (ASCII is a subset of UTF-8)
                _ok(h.H5Tset_cset(mt, h.H5Tget_cset(ft)), "H5Tset_cset")
                if h.H5Tis_variable_str(ft) > 0:
                    _ok(h.H5Tset_size(mt, _T_VARIABLE), "H5Tset_size")
                    buf = (C.c_char_p * n)()
                    _ok(read(mt, C.cast(buf, C.c_void_p)), f"read {what}")
                    try:
                        return [(b or b"").decode("utf-8", "replace") for b in buf]
                    finally:
                        h.H5Dvlen_reclaim(mt, sp, _P_DEFAULT, C.cast(buf, C.c_void_p))
                size = int(h.H5Tget_size(ft))
                _ok(h.H5Tset_size(mt, size), "H5Tset_size")
                raw = C.create_string_buffer(size * n)
                _ok(read(mt, C.cast(raw, C.c_void_p)), f"read {what}")
                b = raw.raw
                return [b[i * size:(i + 1) * size].split(b"\0", 1)[0].decode("utf-8", "replace")
                        for i in range(n)]
            finally:
                h.H5Tclose(mt)
        finally:
            h.H5Tclose(ft)
            h.H5Sclose(sp)

    # ---- attributes -----------------------------------------------------------
    def has_attr(self, path: str, name: str) -> bool:
        oid = self._open_obj(path)
        try:
            return self.h.H5Aexists(oid, name.encode()) > 0
        finally:
            self.h.H5Oclose(oid)

    def attr(self, path: str, name: str):
        """An attribute: str or list of str for string types, a numpy scalar
        or array for numeric ones (enums, such as h5py's booleans, as their
        integer values)."""
        h = self.h
        oid = self._open_obj(path)
        try:
            if h.H5Aexists(oid, name.encode()) <= 0:
                raise KeyError(f"{self.path}:{path} has no attribute {name!r}")
            aid = h.H5Aopen(oid, name.encode(), _P_DEFAULT)
            _ok(aid, f"open attribute {name}")
            try:
                ft = h.H5Aget_type(aid)
                try:
                    cls = h.H5Tget_class(ft)
                finally:
                    h.H5Tclose(ft)
                sp = h.H5Aget_space(aid)
                try:
                    dims = _dims(h, sp)
                finally:
                    h.H5Sclose(sp)
                if cls == _T_STRING:
                    s = self._strings(aid, h.H5Aget_space, h.H5Aget_type,
                                      lambda mt, buf: h.H5Aread(aid, mt, buf), f"{path}@{name}")
                    return s[0] if dims == () else s
                ft = h.H5Aget_type(aid)
                try:
                    dt = _np_dtype(h, ft)
                    if cls == _T_ENUM:
                        # enums convert only to enums: read the file's own bytes
                        mt = h.H5Tget_native_type(ft, _T_DIR_DEFAULT)
                    else:
                        mt = h.H5Tcopy(h.native[dt])
                    try:
                        out = np.empty(dims, dt)
                        _ok(h.H5Aread(aid, mt, out.ctypes.data_as(C.c_void_p)), f"read {name}")
                    finally:
                        h.H5Tclose(mt)
                finally:
                    h.H5Tclose(ft)
                return out[()] if dims == () else out
            finally:
                h.H5Aclose(aid)
        finally:
            h.H5Oclose(oid)

    # ---- writing -------------------------------------------------------------
    def create_group(self, path: str):
        gid = self.h.H5Gcreate2(self.fid, path.encode(), _P_DEFAULT, _P_DEFAULT, _P_DEFAULT)
        _ok(gid, f"create group {path}")
        self.h.H5Gclose(gid)

    def write(self, path: str, a: np.ndarray, chunk: Optional[int] = None,
              compression: Optional[int] = None):
        """A numeric dataset in `a`'s own type; `chunk` elements per chunk
        (1-D) and deflate level `compression` make it chunked and compressed."""
        h = self.h
        a = np.ascontiguousarray(a)
        if a.dtype not in h.native:
            raise H5Error(f"{path}: no native HDF5 type for {a.dtype}")
        dims = (_hsize * max(a.ndim, 1))(*(a.shape or (1,)))
        sp = h.H5Screate_simple(a.ndim, dims, None) if a.ndim else h.H5Screate(_S_SCALAR)
        dcpl = _P_DEFAULT
        try:
            if (chunk or compression is not None) and a.ndim == 1 and a.size:
                dcpl = h.H5Pcreate(h.dcpl_class)
                c = (_hsize * 1)(max(1, min(int(chunk or 65536), a.size)))
                _ok(h.H5Pset_chunk(dcpl, 1, c), "H5Pset_chunk")
                if compression is not None:
                    _ok(h.H5Pset_deflate(dcpl, int(compression)), "H5Pset_deflate")
            did = h.H5Dcreate2(self.fid, path.encode(), h.native[a.dtype], sp, _P_DEFAULT, dcpl,
                               _P_DEFAULT)
            _ok(did, f"create dataset {path}")
            try:
                if a.size:
                    _ok(h.H5Dwrite(did, h.native[a.dtype], _S_ALL, _S_ALL, _P_DEFAULT,
                                   a.ctypes.data_as(C.c_void_p)), f"write {path}")
            finally:
                h.H5Dclose(did)
        finally:
            if dcpl != _P_DEFAULT:
                h.H5Pclose(dcpl)
            h.H5Sclose(sp)

    def _vlen_str_type(self):
        mt = self.h.H5Tcopy(self.h.c_s1)
        _ok(self.h.H5Tset_size(mt, _T_VARIABLE), "H5Tset_size")
        _ok(self.h.H5Tset_cset(mt, _T_CSET_UTF8), "H5Tset_cset")
        return mt

    def _fixed_str_type(self, size: int):
        mt = self.h.H5Tcopy(self.h.c_s1)
        _ok(self.h.H5Tset_size(mt, max(size, 1)), "H5Tset_size")
        _ok(self.h.H5Tset_cset(mt, _T_CSET_UTF8), "H5Tset_cset")
        return mt

    @staticmethod
    def _encode(strings: Sequence[str], fixed: bool):
        enc = [s.encode("utf-8") for s in strings]
        if fixed:
            size = max([len(e) for e in enc] + [1])
            return enc, size, b"".join(e.ljust(size, b"\0") for e in enc)
        return enc, 0, (C.c_char_p * max(len(enc), 1))(*enc)

    def write_strings(self, path: str, strings: Sequence[str], fixed: bool = False):
        """A 1-D string dataset: variable-length UTF-8 (h5py's and anndata's
        form), or fixed-length null-padded when `fixed`."""
        h = self.h
        enc, size, buf = self._encode(strings, fixed)
        mt = self._fixed_str_type(size) if fixed else self._vlen_str_type()
        dims = (_hsize * 1)(len(enc))
        sp = h.H5Screate_simple(1, dims, None)
        try:
            did = h.H5Dcreate2(self.fid, path.encode(), mt, sp, _P_DEFAULT, _P_DEFAULT,
                               _P_DEFAULT)
            _ok(did, f"create dataset {path}")
            try:
                if enc:
                    data = C.c_char_p(buf) if fixed else buf
                    _ok(h.H5Dwrite(did, mt, _S_ALL, _S_ALL, _P_DEFAULT,
                                   C.cast(data, C.c_void_p)), f"write {path}")
            finally:
                h.H5Dclose(did)
        finally:
            h.H5Sclose(sp)
            h.H5Tclose(mt)

    def set_attr(self, path: str, name: str, value):
        """A str (variable-length UTF-8 scalar, as h5py writes it), a list of
        str (1-D, variable length) or a numeric scalar / array attribute."""
        h = self.h
        oid = self._open_obj(path)
        try:
            if isinstance(value, str) or (isinstance(value, (list, tuple)) and value and
                                          all(isinstance(v, str) for v in value)):
                items = [value] if isinstance(value, str) else list(value)
                enc, _, buf = self._encode(items, False)
                mt = self._vlen_str_type()
                sp = h.H5Screate(_S_SCALAR) if isinstance(value, str) else \
                    h.H5Screate_simple(1, (_hsize * 1)(len(items)), None)
                data = C.cast(buf, C.c_void_p)
            else:
                a = np.ascontiguousarray(value)
                if a.dtype == np.bool_:
                    a = a.astype(np.int8)
                mt = h.H5Tcopy(h.native[a.dtype])
                sp = h.H5Screate_simple(a.ndim, (_hsize * a.ndim)(*a.shape), None) if a.ndim \
                    else h.H5Screate(_S_SCALAR)
                data = a.ctypes.data_as(C.c_void_p)
            try:
                aid = h.H5Acreate2(oid, name.encode(), mt, sp, _P_DEFAULT, _P_DEFAULT)
                _ok(aid, f"create attribute {name}")
                try:
                    _ok(h.H5Awrite(aid, mt, data), f"write attribute {name}")
                finally:
                    h.H5Aclose(aid)
            finally:
                h.H5Sclose(sp)
                h.H5Tclose(mt)
        finally:
            h.H5Oclose(oid)


# ---- the harness's loaders (test/test_basic.py:7-58) --------------------------

def _csr_arrays(f: H5File, group: str = "/X"):
    if f.kind(group) != "group" or any(f.kind(f"{group}/{k}") != "dataset"
                                       for k in ("data", "indices", "indptr")):
        raise H5Error(f"{f.path}: {group} is not a CSR group (data, indices, indptr)")
    if f.has_attr(group, "encoding-type"):
        enc = f.attr(group, "encoding-type")
        if enc != "csr_matrix":
            raise H5Error(f"{f.path}: {group} is encoded as {enc!r}, not csr_matrix")
    shape = f.attr(group, "shape")
    return int(shape[0]), int(shape[1])


def _as_float32(a: np.ndarray) -> np.ndarray:
    """Values as the reference's binding narrows them: Python floats (C
    doubles) to C float, round to nearest."""
    if a.dtype == np.float32:
        return a
    return a.astype(np.float64).astype(np.float32)


def _as_int32(a: np.ndarray, what: str) -> np.ndarray:
    if a.dtype == np.int32:
        return a
    if a.size and (a.min() < np.iinfo(np.int32).min or a.max() > np.iinfo(np.int32).max):
        raise H5Error(f"{what}: values outside int32")
    return a.astype(np.int32)


def load_matrix(filename: str):
    """(height, width, data float32, indices int32, indptr int64) of ``/X``
    (test/test_basic.py:7-15: ``/X``'s ``shape`` attribute and its data,
    indices and indptr datasets)."""
    with H5File(filename) as f:
        height, width = _csr_arrays(f)
        data = _as_float32(f.read("/X/data"))
        indices = _as_int32(f.read("/X/indices"), "/X/indices")
        indptr = f.read("/X/indptr").astype(np.int64, copy=False)
    if indptr.shape[0] != height + 1 or (height and indptr[-1] != data.shape[0]) or \
            data.shape[0] != indices.shape[0]:
        raise H5Error(f"{filename}: /X is not a consistent CSR of {height} rows")
    return height, width, data, indices, indptr


def load_celltypes(filename: str) -> Tuple[List[str], np.ndarray]:
    """(label names in code order, per-cell label codes int32) of the
    reference (test/test_basic.py:17-43): the categorical ``/obs/celltype``
    or ``/obs/ClusterName``, else the string dataset
    ``/obsm/annotation_au/celltype`` coded by first appearance."""
    with H5File(filename) as f:
        for g in ("/obs/celltype", "/obs/ClusterName"):
            if g in f:
                if f.kind(g) != "group":
                    raise H5Error(f"{filename}: {g} is not a categorical group "
                                  "(codes, categories)")
                names = f.read_strings(g + "/categories")
                codes = f.read(g + "/codes").astype(np.int64)
                if codes.size and (codes.min() < 0 or codes.max() >= len(names)):
                    raise H5Error(f"{filename}: {g} has cells without a category "
                                  "(code -1) or codes past its categories")
                return names, codes.astype(np.int32)
        g = "/obsm/annotation_au/celltype"
        if g in f:
            vals = f.read_strings(g)
            first: Dict[str, int] = {}
            cells = np.empty(len(vals), np.int32)
            for i, t in enumerate(vals):
                cells[i] = first.setdefault(t, len(first))
            return list(first), cells
    raise H5Error(f"{filename}: no celltype found (/obs/celltype, /obs/ClusterName, "
                  "/obsm/annotation_au/celltype)")


def load_cellnames(filename: str) -> List[str]:
    """The query's cell names, ``/obs/_index`` (test/test_basic.py:45-48)."""
    with H5File(filename) as f:
        col = f.attr("/obs", "_index") if f.has_attr("/obs", "_index") else "_index"
        return f.read_strings(f"/obs/{col}")


def load_geneidx(filename: str) -> List[str]:
    """Gene names: ``/var``'s index column (its ``_index`` attribute), or
    ``Symbol`` when the index column is ``Gene_ID`` (test/test_basic.py:50-56)."""
    with H5File(filename) as f:
        name = f.attr("/var", "_index")
        if name == "Gene_ID":
            name = "Symbol"
        return f.read_strings(f"/var/{name}")


def write_tsv(out_file: str, res) -> None:
    """``cell / firstLabel / finalLabel`` rows as the harness writes them
    (np.savetxt, tab-separated, a header line without comment marker;
    test/test_basic.py:71-73, src/pipeline.cpp:113-121)."""
    with open(out_file, "w") as fh:
        fh.write("cell\tfirstLabel\tfinalLabel\n")
        for row in res:
            fh.write("\t".join(str(x) for x in row) + "\n")


def main(ref_file: str, qry_file: str, out_file: str, cores: int = 20, gpuid: int = 0,
         **run_kwargs) -> list:
    """The harness (test/test_basic.py:58-73): load both files, run, write the
    TSV.  Extra keyword arguments go to ``cusingler.run`` (``path=`` etc.)."""
    from cusingler import cusingler
    ref_height, ref_width, ref_data, ref_indices, ref_indptr = load_matrix(ref_file)
    qry_height, qry_width, qry_data, qry_indices, qry_indptr = load_matrix(qry_file)
    codes, celltypes = load_celltypes(ref_file)
    cellnames = load_cellnames(qry_file)
    ref_geneidx = load_geneidx(ref_file)
    qry_geneidx = load_geneidx(qry_file)
    run_kwargs.setdefault("quantile", 80)
    run_kwargs.setdefault("finetune_thre", 0.05)
    run_kwargs.setdefault("finetune_times", 0)
    res = cusingler.run(cores, gpuid, ref_height, ref_width, ref_data, ref_indices,
                        ref_indptr.astype(np.int32), qry_height, qry_width, qry_data,
                        qry_indices, qry_indptr.astype(np.int32), codes, celltypes, cellnames,
                        ref_geneidx, qry_geneidx, **run_kwargs)
    write_tsv(out_file, res)
    return res


# ---- streaming query batches ---------------------------------------------------

class CsrRows:
    """Row batches of an h5ad's CSR ``/X``: the row pointers are read once;
    each batch's values and column indices are read by hyperslab straight into
    caller buffers (float32 / int32; files in other types are converted as
    ``load_matrix`` converts them)."""

    def __init__(self, filename: str):
        self.f = H5File(filename)
        self.height, self.width = _csr_arrays(self.f)
        self.indptr = self.f.read("/X/indptr").astype(np.int64, copy=False)
        if self.indptr.shape[0] != self.height + 1:
            raise H5Error(f"{filename}: /X/indptr has {self.indptr.shape[0]} entries for "
                          f"{self.height} rows")
        self.nnz = int(self.indptr[-1]) if self.height else 0
        self._data_t = self.f.dtype("/X/data")
        self._ind_t = self.f.dtype("/X/indices")

    def close(self):
        self.f.close()

    def nnz_of(self, r0: int, r1: int) -> int:
        return int(self.indptr[r1] - self.indptr[r0])

    def read(self, r0: int, r1: int, data: np.ndarray, indices: np.ndarray,
             indptr: np.ndarray) -> int:
        """Rows [r0, r1) into the first nnz entries of `data` (float32) and
        `indices` (int32, or uint16 for widths up to 65536: the library
        converts while reading and fails on a column outside uint16) and the
        first r1 - r0 + 1 of `indptr` (int32, from 0).  Returns nnz."""
        b, e = int(self.indptr[r0]), int(self.indptr[r1])
        n = e - b
        if n > data.shape[0] or n > indices.shape[0] or r1 - r0 + 1 > indptr.shape[0]:
            raise ValueError(f"rows [{r0}, {r1}) hold {n} entries; the buffers are too small")
        if n >= 2 ** 31:
            raise ValueError(f"rows [{r0}, {r1}) hold {n} entries (int32 row pointers)")
        if self._data_t == np.float32:
            self.f.read("/X/data", b, e, np.float32, out=data[:n])
        else:
            data[:n] = _as_float32(self.f.read("/X/data", b, e))
        if indices.dtype == np.uint16:
            if self.width > 65536:
                raise ValueError(f"16-bit columns need a width of at most 65536 ({self.width})")
            self.f.read("/X/indices", b, e, np.uint16, out=indices[:n], strict=True)
        elif self._ind_t == np.int32:
            self.f.read("/X/indices", b, e, np.int32, out=indices[:n])
        else:
            indices[:n] = _as_int32(self.f.read("/X/indices", b, e), "/X/indices")
        np.subtract(self.indptr[r0:r1 + 1], b, out=indptr[:r1 - r0 + 1], casting="unsafe")
        return n


def annotate_h5ad(ref_file: str, qry_file: str, out_file: Optional[str] = None,
                  path: str = "plain", batch_rows: int = 262144, quantile: float = 80,
                  finetune_thre: float = 0.05, finetune_times: int = 0, cores: int = 16,
                  gpuid: int = 0, proposal: Optional[Tuple[np.ndarray, int, int, float]] = None,
                  narrow_columns: bool = True, verbose: bool = True) -> Dict[str, object]:
    """Annotates a query .h5ad of any height in batches of `batch_rows` cells.

    The reference is trained on the host (DataParser's steps, one-time) and
    loaded once with room for a batch.  Then, per batch: a host thread reads
    the next batch from the file into one of two page-locked buffer sets; the
    batch after the current one is uploaded on the engine's copy stream
    (csg_prefetch_query_raw) while the current one is value-coded on the
    device for both gene sets (csg_copyin_query_raw), scored and fine-tuned.
    `quantile` is in percent, as ``run`` takes it.  `proposal` =
    (network parameters, hidden width, k, min_margin) turns on the learned
    proposal on the exact path.  `narrow_columns`: batches hold 16-bit
    column indices when the query width allows.

    Returns first / final label indices (int32), the label names, the cell
    names and timings; writes the harness's TSV when `out_file` is given."""
    from . import engine as E
    from .synth import Dataset
    t_all = time.perf_counter()
    rh, rw, rd, ri, rp = load_matrix(ref_file)
    codes, celltype = load_celltypes(ref_file)
    ref_genes = load_geneidx(ref_file)
    qry_genes = load_geneidx(qry_file)
    cellnames = load_cellnames(qry_file)
    rows = CsrRows(qry_file)
    if len(celltype) != rh or len(ref_genes) != rw or len(qry_genes) != rows.width or \
            len(cellnames) != rows.height:
        raise H5Error("the files' obs / var lengths do not match their matrices")
    if rows.height == 0:
        raise H5Error(f"{qry_file}: no query cells")
    if not 0 < batch_rows < 2 ** 31:
        raise ValueError("batch_rows")
    nb = (rows.height + batch_rows - 1) // batch_rows
    bounds = [(k * batch_rows, min(rows.height, (k + 1) * batch_rows)) for k in range(nb)]
    cap_rows = bounds[0][1] - bounds[0][0]
    cap_nnz = max(max(rows.nnz_of(a, b) for a, b in bounds), 1)
    q = np.float32(quantile) / np.float32(100)

    t0 = time.perf_counter()
    # the host steps need a query: the first batch (its coding is not used)
    d0 = np.empty(cap_nnz, np.float32)
    i0 = np.empty(cap_nnz, np.int32)
    p0 = np.empty(cap_rows + 1, np.int32)
    n0 = rows.read(0, cap_rows, d0, i0, p0)
    ds = Dataset(cfg=None, ref_height=rh, ref_width=rw, ref_data=rd, ref_indices=ri,
                 ref_indptr=rp.astype(np.int32), qry_height=cap_rows, qry_width=rows.width,
                 qry_data=d0[:n0], qry_indices=i0[:n0], qry_indptr=p0[:cap_rows + 1].copy(),
                 codes=list(codes), celltype=celltype, cellnames=[], ref_genes=ref_genes,
                 qry_genes=qry_genes, qry_truth=np.zeros(0, np.int32))
    prep = E.Prep(ds, cores)
    c0, c1 = prep.generate(0), prep.generate(1)
    m0, m1 = prep.query_map(0, rows.width), prep.query_map(1, rows.width)
    del ds, d0, i0, p0
    eng = E.Engine(gpuid)
    pin = E.Pinned()
    try:
        eng.reserve(c0, c1, cap_rows)
        dev_path = path
        if proposal is not None:
            params, hidden, k, margin = proposal
            eng.mlp_init(hidden, seed=0)
            eng.set_mlp_params(params)
            eng.set_proposal(k, margin)
            dev_path = "exact_int8"
        del c0, c1
        t_setup = time.perf_counter() - t0
        # columns cross the bus as uint16 when the width allows (6 instead of
        # 8 bytes per stored value)
        it = np.uint16 if (narrow_columns and rows.width <= 65536) else np.int32
        slots = [(pin.empty(cap_nnz, np.float32), pin.empty(cap_nnz, it),
                  pin.empty(cap_rows + 1, np.int32)) for _ in range(2)]
        first = np.empty(rows.height, np.int32)
        final = np.empty(rows.height, np.int32)
        read_s = [0.0]
        err: List[BaseException] = []

        def reader(k: int):
            a, b = bounds[k]
            t = time.perf_counter()
            try:
                rows.read(a, b, *slots[k % 2])
            except BaseException as e:  # reported by the main thread
                err.append(e)
            read_s[0] += time.perf_counter() - t

        def batch(k: int):
            d, i, p = slots[k % 2]
            n = bounds[k][1] - bounds[k][0]
            nnz = int(p[n])
            return d[:nnz], i[:nnz], p[:n + 1]

        t1 = time.perf_counter()
        reader(0)
        if err:
            raise err[0]
        eng.prefetch_query_raw(*batch(0))
        th = None
        if nb > 1:
            th = threading.Thread(target=reader, args=(1,))
            th.start()
        for k in range(nb):
            eng.copyin_query_raw(*batch(k), rows.width, m0, m1)
            if th is not None:
                th.join()
                th = None
                if err:
                    raise err[0]
                eng.prefetch_query_raw(*batch(k + 1))
            f1, _, _ = eng.get_label(float(q), finetune_thre, dev_path, outputs="first")
            f2, _ = eng.finetune(float(q), finetune_thre, finetune_times)
            a, b = bounds[k]
            first[a:b], final[a:b] = f1, f2
            # batch k's buffers are free (its upload finished before the labels)
            if k + 2 < nb:
                th = threading.Thread(target=reader, args=(k + 2,))
                th.start()
        t_stream = time.perf_counter() - t1
    finally:
        eng.close()
        pin.free()
        rows.close()
    if verbose:
        print(f"[h5ad] {rows.height} cells in {nb} batches: set-up {t_setup:.2f} s, "
              f"stream {t_stream:.2f} s ({rows.height / t_stream:,.0f} cells/s; reading "
              f"{read_s[0]:.2f} s on the host thread)", flush=True)
    if out_file is not None:
        write_tsv(out_file, ([c, codes[a], codes[b]] for c, a, b in zip(cellnames, first, final)))
    return {"first": first, "final": final, "codes": list(codes), "cellnames": cellnames,
            "setup_s": t_setup, "stream_s": t_stream, "read_s": read_s[0],
            "total_s": time.perf_counter() - t_all, "batches": nb}


# ---- writing the AnnData layout ---------------------------------------------------

def _dataframe(f: H5File, group: str, index: Sequence[str], index_name: str,
               columns: Dict[str, object], fixed_strings: bool):
    f.create_group(group)
    f.set_attr(group, "_index", index_name)
    f.set_attr(group, "encoding-type", "dataframe")
    f.set_attr(group, "encoding-version", "0.2.0")
    order = list(columns)
    if order:
        f.set_attr(group, "column-order", order)
    else:
        f.set_attr(group, "column-order", np.zeros(0, np.float64))
    f.write_strings(f"{group}/{index_name}", list(index), fixed=fixed_strings)
    f.set_attr(f"{group}/{index_name}", "encoding-type", "string-array")
    f.set_attr(f"{group}/{index_name}", "encoding-version", "0.2.0")
    for name, col in columns.items():
        path = f"{group}/{name}"
        if isinstance(col, tuple):  # categorical: (categories, codes)
            cats, codes = col
            f.create_group(path)
            f.set_attr(path, "encoding-type", "categorical")
            f.set_attr(path, "encoding-version", "0.2.0")
            f.set_attr(path, "ordered", np.bool_(False))
            f.write_strings(f"{path}/categories", list(cats), fixed=fixed_strings)
            f.set_attr(f"{path}/categories", "encoding-type", "string-array")
            f.set_attr(f"{path}/categories", "encoding-version", "0.2.0")
            ct = np.int8 if len(cats) < 128 else (np.int16 if len(cats) < 32768 else np.int32)
            f.write(f"{path}/codes", np.asarray(codes).astype(ct))
            f.set_attr(f"{path}/codes", "encoding-type", "array")
            f.set_attr(f"{path}/codes", "encoding-version", "0.2.0")
        elif isinstance(col, (list, tuple)) or (isinstance(col, np.ndarray) and
                                                col.dtype.kind in "UO"):
            f.write_strings(path, [str(x) for x in col], fixed=fixed_strings)
            f.set_attr(path, "encoding-type", "string-array")
            f.set_attr(path, "encoding-version", "0.2.0")
        else:
            f.write(path, np.asarray(col))
            f.set_attr(path, "encoding-type", "array")
            f.set_attr(path, "encoding-version", "0.2.0")


def write_h5ad(filename: str, height: int, width: int, data: np.ndarray, indices: np.ndarray,
               indptr: np.ndarray, obs_names: Sequence[str], var_names: Sequence[str],
               obs: Optional[Dict[str, object]] = None, var: Optional[Dict[str, object]] = None,
               var_index_name: str = "_index", obsm_strings: Optional[Dict[str, Sequence[str]]] = None,
               chunk: Optional[int] = None, compression: Optional[int] = None,
               fixed_strings: bool = False) -> None:
    """An .h5ad in AnnData's on-disk layout (encoding-type / encoding-version
    attributes): ``/X`` a CSR group with a ``shape`` attribute, ``/obs`` and
    ``/var`` dataframes (string index; a column given as (categories, codes)
    is written as a categorical, a list of str as a string array), empty
    ``/obsm`` (or string datasets under ``/obsm/<group>/<name>`` from
    `obsm_strings`, keys "group/name"), ``/varm``, ``/obsp``, ``/varp``,
    ``/layers`` and ``/uns``.  `chunk` and `compression` (deflate level) apply
    to ``/X``'s datasets; arrays are written in their own dtypes."""
    data, indices, indptr = np.asarray(data), np.asarray(indices), np.asarray(indptr)
    if indptr.shape[0] != height + 1 or data.shape[0] != indices.shape[0] or \
            (height and int(indptr[-1]) != data.shape[0]):
        raise ValueError("inconsistent CSR arrays")
    if len(obs_names) != height or len(var_names) != width:
        raise ValueError("obs / var names do not match the matrix shape")
    with H5File(filename, "w") as f:
        f.set_attr("/", "encoding-type", "anndata")
        f.set_attr("/", "encoding-version", "0.1.0")
        f.create_group("/X")
        f.set_attr("/X", "encoding-type", "csr_matrix")
        f.set_attr("/X", "encoding-version", "0.1.0")
        f.set_attr("/X", "shape", np.array([height, width], np.int64))
        f.write("/X/data", data, chunk, compression)
        f.write("/X/indices", indices, chunk, compression)
        f.write("/X/indptr", indptr, chunk, compression)
        _dataframe(f, "/obs", obs_names, "_index", dict(obs or {}), fixed_strings)
        vcols = dict(var or {})
        if var_index_name != "_index" and var_index_name not in vcols:
            raise ValueError("var_index_name must name a column given in `var`")
        if var_index_name == "_index":
            _dataframe(f, "/var", var_names, "_index", vcols, fixed_strings)
        else:
            # the index is one of the columns (as anndata writes a named index)
            idx = vcols.pop(var_index_name)
            _dataframe(f, "/var", idx, var_index_name, vcols, fixed_strings)
        for g in ("/obsm", "/varm", "/obsp", "/varp", "/layers", "/uns"):
            f.create_group(g)
            f.set_attr(g, "encoding-type", "dict")
            f.set_attr(g, "encoding-version", "0.1.0")
        for key, vals in (obsm_strings or {}).items():
            grp, name = key.split("/", 1)
            if f"/obsm/{grp}" not in f:
                f.create_group(f"/obsm/{grp}")
            f.write_strings(f"/obsm/{grp}/{name}", list(vals), fixed=fixed_strings)


def write_dataset(ds, ref_file: str, qry_file: str, **kw) -> None:
    """A synthetic dataset (``cusingler_amd.synth.Dataset``) as the reference
    and query .h5ad files the harness reads: the reference's labels as the
    categorical ``obs/celltype``, the genes as ``var``'s index."""
    write_h5ad(ref_file, ds.ref_height, ds.ref_width, ds.ref_data, ds.ref_indices, ds.ref_indptr,
               [f"ref_{i}" for i in range(ds.ref_height)], list(ds.ref_genes),
               obs={"celltype": (list(ds.codes), np.asarray(ds.celltype))}, **kw)
    write_h5ad(qry_file, ds.qry_height, ds.qry_width, ds.qry_data, ds.qry_indices, ds.qry_indptr,
               list(ds.cellnames) if len(ds.cellnames) == ds.qry_height else
               [f"cell_{i}" for i in range(ds.qry_height)], list(ds.qry_genes), **kw)


if __name__ == "__main__":
    if len(sys.argv) != 4:
        print("enter <ref file> <qry file> <out stat file>")
        sys.exit(-1)
    main(sys.argv[1], sys.argv[2], sys.argv[3])
