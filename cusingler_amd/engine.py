"""ctypes bindings of ``libcusingler_amd.so`` (include/cusingler_amd.h and
include/cusingler_prep.h) for step-by-step use: the host preprocessing
(``Prep``), the device engine (``Engine``) and a convenience ``annotate`` that
runs train -> score -> finetune like the reference's pipeline
(src/pipeline.cpp:128-164) and returns label indices and scores.

There is no Python fallback: a missing library raises, and the engine
functions need a GPU.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib

PATHS = {"plain": 0, "exact_int8": 1, "fp32_mfma": 2, "fp16": 3}
PHASES = ["rank", "corr", "select", "label", "recheck", "finetune", "ft_corr", "ft_rank",
          "ft_select", "propose", "query_coding"]
COUNTERS = ["cells", "G_s", "ref_cells", "rechecked_pairs", "uncertified_cells", "nan_segments",
            "ft_rounds", "ft_items", "ft_groups", "ft_pairs", "ft_pair_genes", "ft_redone_cells",
            "ft_host_lists_us", "ft_host_group_us", "ft_host_build_us", "ft_batch_us",
            "prop_cells", "prop_fallback_margin", "prop_fallback_check", "prop_scored_pairs",
            "prop_widened", "prop_correlations", "ft_fused_groups", "ft_chained_rounds",
            "ft_chain_halts", "ft_tail_rounds"]
N_PHASES = 11
N_COUNTERS = 26

_V = C.c_void_p
_U32 = C.c_uint32
_I32 = C.c_int32
_U64 = C.c_uint64
_I64 = C.c_int64
_F = C.c_float


class CsgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class CodedStruct(C.Structure):
    _fields_ = [("ref_height", _U32), ("qry_height", _U32), ("width", _U32), ("ct_num", _U32),
                ("ref_data", _V), ("ref_indices", _V), ("ref_indptr", _V),
                ("qry_data", _V), ("qry_indices", _V), ("qry_indptr", _V),
                ("max_uniq_gene", _U64), ("ctidx", _V), ("ctdidx", _V), ("ctdiff", _V),
                ("n_ctdiff", _U64), ("genes", _V)]


_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    L = _lib.load("libcusingler_amd.so")
    P = C.POINTER
    L.csg_init_gpu.argtypes = [C.c_int, P(_V)]
    L.csg_last_error.restype = C.c_char_p
    L.csg_last_error.argtypes = [_V]
    L.csg_release.argtypes = [_V]
    L.csg_copyin_score.argtypes = [_V, _U32, _U32, _V, _V, _V, _U32, _V, _V, _V, _U32, _V, _U64]
    L.csg_get_label.argtypes = [_V, _F, _F, C.c_int, _V, _V, _V]
    L.csg_destroy_score.argtypes = [_V]
    L.csg_copyin.argtypes = [_V, _U32, _U32, _V, _V, _V, _U32, _V, _V, _V, _U32, _V, _V, _V, _U64,
                             _U64, _V]
    L.csg_finetune.argtypes = [_V, _F, _F, C.c_int, _V, _V]
    L.csg_destroy.argtypes = [_V]
    L.csg_set_finetune_path.argtypes = [_V, C.c_int]
    L.csg_set_finetune_schedule.argtypes = [_V, C.c_int]
    L.csg_score_matrix.argtypes = [_V, C.c_int, _U32, _U32, _V]
    L.csg_finetune_tuning.argtypes = [_V, C.c_int, _V, _V]
    L.csg_set_finetune_budget.argtypes = [_V, _U64, _V]
    L.csg_set_finetune_fuse.argtypes = [_V, C.c_int]
    L.csg_set_finetune_chain.argtypes = [_V, C.c_int, _U64]
    L.csg_set_finetune_tail.argtypes = [_V, _U32, C.c_int]
    L.csg_set_limb_form.argtypes = [_V, C.c_int, _V]
    L.csg_finetune_tail_profile.argtypes = [_V, _V, C.c_int, _V]
    L.csg_set_query_chunk.argtypes = [_V, _U32, _V]
    L.csg_copyin_query_raw.argtypes = [_V, _U32, _U32, _V, _V, _V, _V, _V]
    L.csg_copyin_query_raw_u16.argtypes = [_V, _U32, _U32, _V, _V, _V, _V, _V]
    L.csg_prefetch_query_raw_u16.argtypes = [_V, _U32, _V, _V, _V]
    L.csg_prep_query_map.argtypes = [_V, C.c_int, _V]
    L.csg_query_coded.argtypes = [_V, C.c_int, _V, _V, _V]
    L.csg_set_upload_slice.argtypes = [_V, _U64]
    L.csg_code_csr.argtypes = [_V, _U32, _U32] + [_V] * 12
    L.csg_code_csr_noalloc.argtypes = [_V, _U32, _U32] + [_V] * 12
    L.csg_synchronize.argtypes = [_V]
    L.csg_host_alloc.argtypes = [_U64, P(_V)]
    L.csg_host_free.argtypes = [_V]
    L.csg_prefetch_query_raw.argtypes = [_V, _U32, _V, _V, _V]
    L.csg_train.argtypes = [_V, _U32, _U32, _V, _V, _V, _V, _U32, _U32, _V, _V]
    L.csg_train_table.restype = _U64
    L.csg_train_table.argtypes = [_V, C.c_int, _V]
    L.csg_prep_ref_map.argtypes = [_V, _V]
    L.csg_phase_times.argtypes = [_V, _V, _V, C.c_int]
    L.csg_set_phase_timing.argtypes = [_V, C.c_int]
    L.csg_counters.argtypes = [_V, _V]
    L.csg_selftest.argtypes = [_V, C.c_int, P(C.c_double)]
    L.csg_selftest_ft.argtypes = [_V, C.c_int, P(_I64)]
    L.csg_host_percentile.argtypes = [_V, _I64, _F, P(_F)]
    L.csg_host_minmax.argtypes = [_V, C.c_int, P(C.c_int), P(C.c_int)]
    L.csg_mlp_init.argtypes = [_V, C.c_int, _U64]
    L.csg_mlp_dims.argtypes = [_V, _V]
    L.csg_mlp_get_params.argtypes = [_V, _V]
    L.csg_mlp_set_params.argtypes = [_V, _V]
    L.csg_mlp_grad.argtypes = [_V, _U32, _V, _V, _V, _V]
    L.csg_mlp_train.argtypes = [_V, _V, C.c_int, C.c_int, _F, _U64, _V]
    L.csg_mlp_predict.argtypes = [_V, C.c_int, _V, _V]
    L.csg_set_proposal.argtypes = [_V, C.c_int, _F]
    L.csg_prep_train.argtypes = [C.c_int, _U32, _U32, _V, _V, _V, _U32, _U32, _V, _V, _V, _U32, _V,
                                 _V, _U32, _V, _U32, P(_V), C.c_char_p, C.c_size_t]
    L.csg_prep_generate.argtypes = [_V, C.c_int, P(CodedStruct)]
    L.csg_prep_train_table.restype = _U64
    L.csg_prep_train_table.argtypes = [_V, C.c_int, _V]
    L.csg_prep_filtered.argtypes = [_V]
    L.csg_prep_release.argtypes = [_V]
    _LIB = L
    return L


def _p(a: np.ndarray) -> _V:
    return _V(a.ctypes.data)


def _np(ptr, n, dtype) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).copy()


def _names(lst: Sequence[str]):
    enc = [s.encode() for s in lst]
    arr = (C.c_char_p * len(enc))(*enc)
    return arr, enc


@dataclass
class Coded:
    """Value-coded CSR matrices of one step (generateDenseMatrix output)."""
    ref_height: int
    qry_height: int
    width: int
    ct_num: int
    ref_data: np.ndarray
    ref_indices: np.ndarray
    ref_indptr: np.ndarray
    qry_data: np.ndarray
    qry_indices: np.ndarray
    qry_indptr: np.ndarray
    max_uniq_gene: int
    ctidx: np.ndarray
    ctdidx: np.ndarray
    ctdiff: np.ndarray
    genes: np.ndarray


class Prep:
    """Host preprocessing (DataParser, src/io.cpp)."""

    def __init__(self, ds, cores: int = 8):
        L = lib()
        self._h = _V()
        rip = np.ascontiguousarray(ds.ref_indptr, dtype=np.int32)
        qip = np.ascontiguousarray(ds.qry_indptr, dtype=np.int32)
        rd = np.ascontiguousarray(ds.ref_data, dtype=np.float32)
        ri = np.ascontiguousarray(ds.ref_indices, dtype=np.int32)
        qd = np.ascontiguousarray(ds.qry_data, dtype=np.float32)
        qi = np.ascontiguousarray(ds.qry_indices, dtype=np.int32)
        ct = np.ascontiguousarray(ds.celltype, dtype=np.int32)
        rg, _k1 = _names(ds.ref_genes)
        qg, _k2 = _names(ds.qry_genes)
        err = C.create_string_buffer(512)
        t0 = time.perf_counter()
        rc = L.csg_prep_train(cores, ds.ref_height, ds.ref_width, _p(rd), _p(ri), _p(rip),
                              ds.qry_height, ds.qry_width, _p(qd), _p(qi), _p(qip), len(ds.codes),
                              _p(ct), C.cast(rg, _V), len(ds.ref_genes), C.cast(qg, _V),
                              len(ds.qry_genes), C.byref(self._h), err, 512)
        self.train_seconds = time.perf_counter() - t0
        if rc != 0:
            raise CsgError(rc, err.value.decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().csg_prep_release(self._h)
            self._h = None

    def generate(self, step: int) -> Coded:
        c = CodedStruct()
        rc = lib().csg_prep_generate(self._h, step, C.byref(c))
        if rc != 0:
            raise CsgError(rc, f"generateDenseMatrix({step}) failed")
        rn = _np(c.ref_indptr, c.ref_height + 1, np.int32)
        qn = _np(c.qry_indptr, c.qry_height + 1, np.int32)
        L = c.ct_num
        nd = 2 * L * L
        return Coded(c.ref_height, c.qry_height, c.width, c.ct_num,
                     _np(c.ref_data, int(rn[-1]), np.uint16), _np(c.ref_indices, int(rn[-1]), np.int32),
                     rn, _np(c.qry_data, int(qn[-1]), np.uint16),
                     _np(c.qry_indices, int(qn[-1]), np.int32), qn, int(c.max_uniq_gene),
                     _np(c.ctidx, 2 * L, np.uint32), _np(c.ctdidx, nd, np.uint32),
                     _np(c.ctdiff, int(c.n_ctdiff), np.uint32), _np(c.genes, c.width, np.uint32))

    def table(self, which: int) -> np.ndarray:
        n = lib().csg_prep_train_table(self._h, which, None)
        out = np.zeros(n, dtype=np.uint32)
        lib().csg_prep_train_table(self._h, which, _p(out))
        return out

    @property
    def filtered(self) -> bool:
        return bool(lib().csg_prep_filtered(self._h))

    def ref_map(self, ref_width: int) -> np.ndarray:
        """Per reference column, its shared column after removeCols or -1."""
        out = np.zeros(ref_width, np.int32)
        rc = lib().csg_prep_ref_map(self._h, _p(out))
        if rc != 0:
            raise CsgError(rc, "ref_map")
        return out

    def query_map(self, step: int, qry_width: int) -> np.ndarray:
        """Per query column (as given to the training call), its column in
        step `step`'s gene set or -1 (the device value coder's map)."""
        out = np.zeros(qry_width, np.int32)
        rc = lib().csg_prep_query_map(self._h, step, _p(out))
        if rc != 0:
            raise CsgError(rc, "query_map")
        return out


class Pinned:
    """Page-locked host arrays from the library's HIP runtime (csg_host_alloc):
    query batches read into them upload by DMA without a staging copy.  The
    arrays are views of that memory; do not use them after free()."""

    def __init__(self):
        self._ptrs: List[_V] = []

    def empty(self, n: int, dtype) -> np.ndarray:
        dt = np.dtype(dtype)
        nbytes = max(int(n) * dt.itemsize, 1)
        p = _V()
        rc = lib().csg_host_alloc(nbytes, C.byref(p))
        if rc != 0 or not p.value:
            raise CsgError(rc or 3, f"csg_host_alloc({nbytes}) failed")
        self._ptrs.append(p)
        buf = (C.c_char * nbytes).from_address(p.value)
        return np.frombuffer(buf, dtype=dt, count=int(n))

    def copy(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a)
        out = self.empty(a.size, a.dtype).reshape(a.shape)
        out[...] = a
        return out

    def free(self):
        for p in self._ptrs:
            lib().csg_host_free(p)
        self._ptrs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class Engine:
    """Device engine (include/cusingler_amd.h)."""

    def __init__(self, gpuid: int = 0):
        self._h = _V()
        rc = lib().csg_init_gpu(gpuid, C.byref(self._h))
        if rc != 0:
            raise CsgError(rc, f"initGPU({gpuid}) failed")
        self._keep: List[np.ndarray] = []

    def close(self):
        if getattr(self, "_h", None):
            lib().csg_release(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise CsgError(rc, f"{what}: {lib().csg_last_error(self._h).decode()}")

    def copyin_score(self, c: Coded):
        self._check(lib().csg_copyin_score(
            self._h, c.ref_height, c.width, _p(c.ref_data), _p(c.ref_indices), _p(c.ref_indptr),
            c.qry_height, _p(c.qry_data), _p(c.qry_indices), _p(c.qry_indptr), c.ct_num,
            _p(c.ctidx), c.max_uniq_gene), "copyin_score")
        self.Nq, self.L, self.Nr = c.qry_height, c.ct_num, c.ref_height

    def reserve(self, c0: Coded, c1: Coded, qry_height: int):
        """Loads the reference sides of both steps (copyin_score and copyin
        without query data) with room for query batches of qry_height rows."""
        self._check(lib().csg_copyin_score(
            self._h, c0.ref_height, c0.width, _p(c0.ref_data), _p(c0.ref_indices),
            _p(c0.ref_indptr), qry_height, None, None, None, c0.ct_num, _p(c0.ctidx),
            c0.max_uniq_gene), "copyin_score (reserve)")
        self.Nq, self.L, self.Nr = qry_height, c0.ct_num, c0.ref_height
        self._check(lib().csg_copyin(
            self._h, c1.ref_height, c1.width, _p(c1.ref_data), _p(c1.ref_indices),
            _p(c1.ref_indptr), qry_height, None, None, None, c1.ct_num, _p(c1.ctidx),
            _p(c1.ctdidx), _p(c1.ctdiff), len(c1.ctdiff), c1.max_uniq_gene, None),
            "copyin (reserve)")
        self.fNq = qry_height

    def copyin_query_raw(self, data: np.ndarray, indices: np.ndarray, indptr: np.ndarray,
                         qry_width: int, map_score: np.ndarray, map_finetune: np.ndarray):
        """A batch of float CSR query rows, value-coded on the device for both
        steps (replaces the loaded query).  Column indices in uint16 (query
        widths up to 65536) take the 16-bit entry point; others are int32."""
        d = np.ascontiguousarray(data, np.float32)
        u16 = isinstance(indices, np.ndarray) and indices.dtype == np.uint16
        i = np.ascontiguousarray(indices, np.uint16 if u16 else np.int32)
        p = np.ascontiguousarray(indptr, np.int32)
        ms = np.ascontiguousarray(map_score, np.int32)
        mf = np.ascontiguousarray(map_finetune, np.int32)
        n = p.shape[0] - 1
        f = lib().csg_copyin_query_raw_u16 if u16 else lib().csg_copyin_query_raw
        self._check(f(self._h, n, qry_width, _p(d), _p(i), _p(p), _p(ms), _p(mf)),
                    "copyin_query_raw")
        self.Nq = self.fNq = n

    def train(self, ds, ref_map: Optional[np.ndarray] = None) -> Dict[str, np.ndarray]:
        """trainData on the device for the reference of dataset `ds`: returns
        train_idxs, train_values, common_genes, thre_genes."""
        d = np.ascontiguousarray(ds.ref_data, np.float32)
        i = np.ascontiguousarray(ds.ref_indices, np.int32)
        p = np.ascontiguousarray(ds.ref_indptr, np.int32)
        ct = np.ascontiguousarray(ds.celltype, np.int32)
        m = None if ref_map is None else np.ascontiguousarray(ref_map, np.int32)
        sizes = np.zeros(3, np.uint64)
        self._check(lib().csg_train(self._h, ds.ref_height, ds.ref_width, _p(d), _p(i), _p(p),
                                    _p(m) if m is not None else None, ds.ref_width,
                                    len(ds.codes), _p(ct), _p(sizes)), "train")
        out = {}
        for k, name in enumerate(["train_idxs", "train_values", "common_genes", "thre_genes"]):
            n = int(lib().csg_train_table(self._h, k, None))
            a = np.zeros(max(n, 1), np.uint32)
            lib().csg_train_table(self._h, k, _p(a))
            out[name] = a[:n]
        return out

    def prefetch_query_raw(self, data: np.ndarray, indices: np.ndarray, indptr: np.ndarray):
        """Starts the upload of the next batch (the arrays must stay alive and
        unchanged until copyin_query_raw consumes them; pass the same arrays)."""
        u16 = indices.dtype == np.uint16
        for a, dt in ((data, np.float32), (indices, np.uint16 if u16 else np.int32),
                      (indptr, np.int32)):
            if a.dtype != dt or not a.flags.c_contiguous:
                raise ValueError("prefetch needs contiguous float32 / int32 (or uint16 column) "
                                 "arrays (the same objects copyin_query_raw will get)")
        f = lib().csg_prefetch_query_raw_u16 if u16 else lib().csg_prefetch_query_raw
        self._check(f(self._h, indptr.shape[0] - 1, _p(data), _p(indices), _p(indptr)),
                    "prefetch_query_raw")

    def query_coded(self, step: int):
        """The query's value-coded CSR of a step as it sits on the device:
        (data uint16, indices int32, indptr int32)."""
        n = self.Nq if step == 0 else self.fNq
        ptr = np.zeros(n + 1, np.int32)
        self._check(lib().csg_query_coded(self._h, step, None, None, _p(ptr)), "query_coded")
        nnz = int(ptr[-1] - ptr[0])
        data = np.zeros(max(nnz, 1), np.uint16)
        ind = np.zeros(max(nnz, 1), np.int32)
        self._check(lib().csg_query_coded(self._h, step, _p(data), _p(ind), _p(ptr)),
                    "query_coded")
        return data[:nnz], ind[:nnz], ptr

    def code_csr(self, rows: int, width: int, data, indices, indptr, map_a, map_b, out_a,
                 out_b):
        """Value coding of a device CSR (device pointers as ints): out_x =
        (data, indices, indptr) device pointers of set x; returns max_uniq."""
        mu = np.zeros(2, np.uint32)
        self._check(lib().csg_code_csr(self._h, rows, width, data, indices, indptr, map_a, map_b,
                                       out_a[0], out_a[1], out_a[2], out_b[0], out_b[1],
                                       out_b[2], _p(mu)), "code_csr")
        return int(mu[0]), int(mu[1])

    def code_csr_noalloc(self, rows: int, width: int, data, indices, indptr, map_a, map_b,
                         out_a, out_b, status):
        """csg_code_csr_noalloc: the same coding with the library's fixed work
        memory; `status` is a device pointer to 4 uint32 (flags, max_uniq a,
        max_uniq b, 0), valid after the engine's stream passes the call."""
        self._check(lib().csg_code_csr_noalloc(self._h, rows, width, data, indices, indptr, map_a,
                                               map_b, out_a[0], out_a[1], out_a[2], out_b[0],
                                               out_b[1], out_b[2], status), "code_csr_noalloc")

    def synchronize(self):
        """Waits for every call made on the engine's stream."""
        self._check(lib().csg_synchronize(self._h), "synchronize")

    def set_upload_slice(self, entries: int):
        self._check(lib().csg_set_upload_slice(self._h, int(entries)), "set_upload_slice")

    def get_label(self, quantile: float = 0.8, thre: float = 0.05, path: str = "plain",
                  outputs=True):
        """outputs: True (first labels, mask and scores to the host), "first"
        (first labels only) or False (results stay on the device)."""
        first = np.zeros(self.Nq, np.int32) if outputs else None
        full = outputs is True
        mask = np.zeros((self.Nq, self.L), np.float32) if full else None
        scores = np.zeros((self.Nq, self.L), np.float32) if full else None
        self._check(lib().csg_get_label(
            self._h, quantile, thre, PATHS[path], _p(first) if outputs else None,
            _p(mask) if full else None, _p(scores) if full else None), "get_label")
        return first, mask, scores

    def destroy_score(self):
        self._check(lib().csg_destroy_score(self._h), "destroy_score")

    def copyin(self, c: Coded, mask: Optional[np.ndarray] = None):
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        self._keep = [m] if m is not None else []
        self._check(lib().csg_copyin(
            self._h, c.ref_height, c.width, _p(c.ref_data), _p(c.ref_indices), _p(c.ref_indptr),
            c.qry_height, _p(c.qry_data), _p(c.qry_indices), _p(c.qry_indptr), c.ct_num,
            _p(c.ctidx), _p(c.ctdidx), _p(c.ctdiff), len(c.ctdiff), c.max_uniq_gene,
            _p(m) if m is not None else None), "copyin")
        self.fNq = c.qry_height

    def set_finetune_path(self, path: str):
        self._check(lib().csg_set_finetune_path(self._h, PATHS[path]), "set_finetune_path")

    def score_matrix(self, path: str, q0: int, nq: int) -> np.ndarray:
        """Per-pair correlations of query rows [q0, q0 + nq) against every
        reference cell by `path`'s contraction (diagnostic)."""
        out = np.zeros((nq, self.Nr), np.float32)
        self._check(lib().csg_score_matrix(self._h, PATHS[path], q0, nq, _p(out)), "score_matrix")
        return out

    def set_finetune_schedule(self, schedule: str):
        """"device" (rounds on the device, HIP graphs) or "host" (host grouping)."""
        self._check(lib().csg_set_finetune_schedule(self._h, {"device": 0, "host": 1}[schedule]),
                    "set_finetune_schedule")

    def set_finetune_budget(self, nbytes: int = 0) -> int:
        """Device bytes of one fine-tune work batch (0 keeps it); returns the
        number of range batches run so far."""
        b = C.c_int64()
        self._check(lib().csg_set_finetune_budget(self._h, nbytes, C.byref(b)), "set_finetune_budget")
        return b.value

    def set_query_chunk(self, rows: int = 0) -> int:
        """Caps the score step's query chunk (a multiple of 128 rows, 0 = no
        cap) from the next copyin_score; returns the current chunk."""
        c = C.c_uint32()
        self._check(lib().csg_set_query_chunk(self._h, int(rows), C.byref(c)), "set_query_chunk")
        return c.value

    def set_finetune_fuse(self, max_cells: int):
        """Largest fine-tune group (cells) of the fused FP32 form (0 = off)."""
        self._check(lib().csg_set_finetune_fuse(self._h, int(max_cells)), "set_finetune_fuse")

    def set_finetune_chain(self, max_rounds: int, max_scores: int = 0):
        """Longest chain of fine-tune rounds launched without a host read-back
        (0 = every round reads its counts back); a chained round with more than
        max_scores correlations (0 = no limit) halts the chain and runs unchained."""
        self._check(lib().csg_set_finetune_chain(self._h, int(max_rounds), int(max_scores)),
                    "set_finetune_chain")

    def set_finetune_tail(self, max_cells: int, blocks_per_cu: int = 4):
        """Rounds with at most max_cells active cells (and all later ones) run in
        one cooperative launch (0 = never)."""
        self._check(lib().csg_set_finetune_tail(self._h, int(max_cells), int(blocks_per_cu)),
                    "set_finetune_tail")

    def set_limb_form(self, form: int) -> int:
        """Exact INT8 contraction's limb form: 0 = the three-product form for
        gene sets of up to 1008 genes (default), 1 = the folded four-product
        form everywhere.  Returns the previous form."""
        prev = C.c_int()
        self._check(lib().csg_set_limb_form(self._h, int(form), C.byref(prev)), "set_limb_form")
        return prev.value

    def finetune_tail_profile(self) -> np.ndarray:
        """Microseconds between the last tail kernel's grid barriers (the first
        round's begin, then per round: rows, end + next begin)."""
        out = np.zeros(64, np.float64)
        n = C.c_int()
        self._check(lib().csg_finetune_tail_profile(self._h, _p(out), 64, C.byref(n)),
                    "finetune_tail_profile")
        if n.value < 0:
            raise CsgError(0, f"the tail kernel was refused (reason {-n.value})")
        return out[:n.value]

    def finetune_tuning(self, autotune: Optional[bool] = None) -> Dict[str, object]:
        """Autotuner state of the device-resident fine-tune (optionally switching
        it on / off, which asks for a new tuning)."""
        g = np.zeros(6, np.int32)
        ms = np.zeros(4, np.float32)
        a = -1 if autotune is None else int(bool(autotune))
        self._check(lib().csg_finetune_tuning(self._h, a, _p(g), _p(ms)), "finetune_tuning")
        return {"tuned": bool(g[0]), "rank_blocks": int(g[1]), "corr_blocks": int(g[2]),
                "select_blocks": int(g[3]), "big_tile_min_cells": int(g[4]),
                "fused_blocks": int(g[5]), "best_ms": [float(v) for v in ms]}

    def finetune(self, quantile: float = 0.8, thre: float = 0.05, times: int = 0):
        final = np.zeros(self.fNq, np.uint32)
        rounds = np.zeros(self.fNq, np.int32)
        self._check(lib().csg_finetune(self._h, quantile, thre, times, _p(final), _p(rounds)),
                    "finetune")
        return final, rounds

    def destroy(self):
        self._check(lib().csg_destroy(self._h), "destroy")

    # ---- learned proposal network ----

    def mlp_init(self, hidden: int = 512, seed: int = 0):
        self._check(lib().csg_mlp_init(self._h, hidden, seed), "mlp_init")

    def mlp_dims(self) -> Dict[str, int]:
        d = np.zeros(6, np.int64)
        self._check(lib().csg_mlp_dims(self._h, _p(d)), "mlp_dims")
        return dict(zip(["G", "Kp", "H", "L", "Lp", "n_params"], (int(v) for v in d)))

    def mlp_params(self) -> np.ndarray:
        w = np.zeros(self.mlp_dims()["n_params"], np.float32)
        self._check(lib().csg_mlp_get_params(self._h, _p(w)), "mlp_get_params")
        return w

    def set_mlp_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.shape != (self.mlp_dims()["n_params"],):
            raise ValueError("parameter array has the wrong size")
        self._check(lib().csg_mlp_set_params(self._h, _p(w)), "mlp_set_params")

    def mlp_grad(self, x: np.ndarray, y: np.ndarray):
        """Loss and gradient (flat, parameter layout) of the mean cross-entropy
        over the rows of x ([n x G_s] features) with labels y."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        g = np.zeros(self.mlp_dims()["n_params"], np.float32)
        loss = C.c_float()
        self._check(lib().csg_mlp_grad(self._h, x.shape[0], _p(x), _p(y), _p(g), C.byref(loss)),
                    "mlp_grad")
        return g, loss.value

    def mlp_train(self, labels: np.ndarray, epochs: int = 4, batch: int = 1024, lr: float = 1e-3,
                  seed: int = 0) -> np.ndarray:
        y = np.ascontiguousarray(labels, dtype=np.int32)
        if y.shape != (self.Nq,):
            raise ValueError("one label per loaded query cell required")
        losses = np.zeros(epochs, np.float32)
        self._check(lib().csg_mlp_train(self._h, _p(y), epochs, batch, lr, seed, _p(losses)),
                    "mlp_train")
        return losses

    def mlp_predict(self, k: int = 4):
        top = np.zeros((self.Nq, k), np.int32)
        margin = np.zeros(self.Nq, np.float32)
        self._check(lib().csg_mlp_predict(self._h, k, _p(top), _p(margin)), "mlp_predict")
        return top, margin

    def set_proposal(self, k: int, min_margin: float = 0.2):
        """Proposal-restricted exact scoring (k labels per cell; 0 turns it off)."""
        self._check(lib().csg_set_proposal(self._h, k, min_margin), "set_proposal")

    def set_phase_timing(self, mode: int) -> None:
        """0: no phase events, 1: the correlation phase only, 2: every phase."""
        self._check(lib().csg_set_phase_timing(self._h, int(mode)), "set_phase_timing")

    def phase_times(self, reset: bool = False) -> Dict[str, tuple]:
        ms = np.zeros(N_PHASES, np.float64)
        cnt = np.zeros(N_PHASES, np.int64)
        self._check(lib().csg_phase_times(self._h, _p(ms), _p(cnt), int(reset)), "phase_times")
        return {PHASES[i]: (float(ms[i]), int(cnt[i])) for i in range(N_PHASES)}

    def selftest(self, which: int) -> float:
        out = C.c_double()
        self._check(lib().csg_selftest(self._h, which, C.byref(out)), "selftest")
        return out.value

    def selftest_ft(self, gs: int) -> int:
        out = _I64()
        self._check(lib().csg_selftest_ft(self._h, gs, C.byref(out)), "selftest_ft")
        return out.value

    def counters(self) -> Dict[str, int]:
        out = np.zeros(N_COUNTERS, np.int64)
        self._check(lib().csg_counters(self._h, _p(out)), "counters")
        return {COUNTERS[i]: int(out[i]) for i in range(N_COUNTERS)}


def certify_width(G: int) -> float:
    """Half-width within which the exact path's per-label quantiles bound the
    reference's FP32 ones (the engine's certification width)."""
    u = 2.0 ** -24
    n = max(G - 1, 0)
    gamma = n * u / (1 - n * u)
    ed = gamma + 2 * u * (1 + gamma)
    wpair = (gamma + ed) / (1 - ed) + 2 * u
    return wpair * (1 + 1e-6) + 16 * u


def host_percentile(arr: np.ndarray, quantile: float) -> float:
    a = np.ascontiguousarray(arr, dtype=np.float32)
    out = C.c_float()
    rc = lib().csg_host_percentile(_p(a), a.shape[0], quantile, C.byref(out))
    if rc != 0:
        raise CsgError(rc, "percentile undefined for this input")
    return out.value


def host_minmax(arr: np.ndarray):
    a = np.ascontiguousarray(arr, dtype=np.float32)
    mn, mx = C.c_int(), C.c_int()
    lib().csg_host_minmax(_p(a), a.shape[0], C.byref(mn), C.byref(mx))
    return mn.value, mx.value


@dataclass
class Annotation:
    first: np.ndarray
    final: np.ndarray
    mask: np.ndarray
    scores: np.ndarray
    rounds: np.ndarray
    timings: Dict[str, float] = field(default_factory=dict)
    counters: Dict[str, int] = field(default_factory=dict)


def annotate(ds, path: str = "plain", quantile: float = 0.8, thre: float = 0.05, times: int = 0,
             cores: int = 8, gpuid: int = 0, prep: Optional[Prep] = None,
             engine: Optional[Engine] = None) -> Annotation:
    """train -> score -> finetune on one dataset; label indices are returned."""
    prep = prep or Prep(ds, cores)
    eng = engine or Engine(gpuid)
    c0 = prep.generate(0)
    eng.copyin_score(c0)
    t0 = time.perf_counter()
    first, mask, scores = eng.get_label(quantile, thre, path)
    t1 = time.perf_counter()
    cnt = eng.counters()
    eng.destroy_score()
    c1 = prep.generate(1)
    eng.copyin(c1)
    t2 = time.perf_counter()
    final, rounds = eng.finetune(quantile, thre, times)
    t3 = time.perf_counter()
    cnt2 = eng.counters()
    cnt.update({k: v for k, v in cnt2.items() if k.startswith("ft_")})
    eng.destroy()
    return Annotation(first, final, mask, scores, rounds,
                      {"score_s": t1 - t0, "finetune_s": t3 - t2, "train_s": prep.train_seconds},
                      cnt)
