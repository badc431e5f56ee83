"""ctypes bindings of the CPU oracle (oracle/liboracle_singler.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_singler.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_V = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int
_F = C.c_float

VAR_SEQ = 0      # FP32, score-step sequential order
VAR_EXACT = 1    # exact integer sums, FP64 normalisation, rounded to float
VAR_TREE = 2     # FP32, fine-tune 128-chunk + tree order

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_train.restype = _V
        L.or_train.argtypes = [_I, _U32, _U32, _V, _V, _V, _U32, _U32, _V, _V, _V, _I, _V, _V, _I,
                               _V, _I]
        L.or_last_error.restype = C.c_char_p
        L.or_free.argtypes = [_V]
        L.or_sizes.argtypes = [_V, _V]
        L.or_get.argtypes = [_V, _I, _V]
        L.or_rank.argtypes = [_V, _U32, _V]
        L.or_percentile.restype = _F
        L.or_percentile.argtypes = [_V, _I, _F]
        L.or_minmax.argtypes = [_V, _I, C.POINTER(_I), C.POINTER(_I)]
        L.or_corr.restype = _F
        L.or_corr.argtypes = [_V, _V, _U32, _I, C.POINTER(C.c_double)]
        L.or_score_matrix.argtypes = [_V, _U32, _U32, _I, _V, _V]
        L.or_score.argtypes = [_V, _U32, _U32, _F, _F, _I, _V, _V, _V]
        L.or_finetune.argtypes = [_V, _U32, _U32, _V, _F, _F, _I, _I, _V, _V]
        L.or_finetune_round.restype = _I
        L.or_finetune_round.argtypes = [_V, _U32, _V, _I, _F, _F, _I, _V, _V]
        _lib = L
    return _lib


def _p(a):
    return _V(a.ctypes.data)


def _blob(names):
    b = ("\0".join(names) + "\0").encode()
    return C.create_string_buffer(b, len(b))


SIZE_KEYS = ["L", "Gs", "Gf", "max_uniq0", "max_uniq1", "n_train_values", "n_ctdiff1",
             "ref_nnz0", "qry_nnz0", "ref_nnz1", "qry_nnz1", "Nr", "Nq"]


class Oracle:
    """The CPU restatement of the reference's pipeline on one dataset."""

    def __init__(self, ds, cores: int = 8):
        L = lib()
        self._keep = [np.ascontiguousarray(ds.ref_data, np.float32),
                      np.ascontiguousarray(ds.ref_indices, np.int32),
                      np.ascontiguousarray(ds.ref_indptr, np.int32),
                      np.ascontiguousarray(ds.qry_data, np.float32),
                      np.ascontiguousarray(ds.qry_indices, np.int32),
                      np.ascontiguousarray(ds.qry_indptr, np.int32),
                      np.ascontiguousarray(ds.celltype, np.int32)]
        rd, ri, rp, qd, qi, qp, ct = self._keep
        self._rn = _blob(ds.ref_genes)
        self._qn = _blob(ds.qry_genes)
        h = L.or_train(cores, ds.ref_height, ds.ref_width, _p(rd), _p(ri), _p(rp), ds.qry_height,
                       ds.qry_width, _p(qd), _p(qi), _p(qp), len(ds.codes), _p(ct), self._rn,
                       len(ds.ref_genes), self._qn, len(ds.qry_genes))
        if not h:
            raise RuntimeError(L.or_last_error().decode())
        self._h = _V(h)
        s = (C.c_int64 * 13)()
        L.or_sizes(self._h, s)
        self.sizes = dict(zip(SIZE_KEYS, list(s)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_free(self._h)
            self._h = None

    def get(self, which: int, n: int, dtype) -> np.ndarray:
        out = np.zeros(max(n, 1), dtype=dtype)
        lib().or_get(self._h, which, _p(out))
        return out[:n]

    def coded(self, step: int):
        s = self.sizes
        base = 10 if step == 0 else 20
        rn, qn = (s["ref_nnz0"], s["qry_nnz0"]) if step == 0 else (s["ref_nnz1"], s["qry_nnz1"])
        return dict(
            ref_data=self.get(base, rn, np.uint16), ref_indices=self.get(base + 1, rn, np.int32),
            ref_indptr=self.get(base + 2, s["Nr"] + 1, np.int32),
            qry_data=self.get(base + 3, qn, np.uint16), qry_indices=self.get(base + 4, qn, np.int32),
            qry_indptr=self.get(base + 5, s["Nq"] + 1, np.int32))

    def tables(self):
        s = self.sizes
        L = s["L"]
        return dict(ctidx=self.get(0, 2 * L, np.uint32),
                    train_idxs=self.get(1, 2 * L * L, np.uint32),
                    train_values=self.get(2, s["n_train_values"], np.uint32),
                    common_genes=self.get(3, s["Gs"], np.uint32),
                    thre_genes=self.get(4, s["Gf"], np.uint32),
                    ctdidx1=self.get(5, 2 * L * L, np.uint32),
                    ctdiff1=self.get(6, s["n_ctdiff1"], np.uint32))

    def score_matrix(self, q0: int, q1: int, variant: int = VAR_SEQ, want64: bool = False):
        Nr = self.sizes["Nr"]
        out = np.zeros((q1 - q0, Nr), np.float32)
        out64 = np.zeros((q1 - q0, Nr), np.float64) if want64 else None
        lib().or_score_matrix(self._h, q0, q1, variant, _p(out), _p(out64) if want64 else None)
        return (out, out64) if want64 else out

    def score(self, q0: int, q1: int, quantile=0.8, thre=0.05, variant: int = VAR_SEQ):
        L = self.sizes["L"]
        n = q1 - q0
        first = np.zeros(n, np.int32)
        mask = np.zeros((n, L), np.float32)
        scores = np.zeros((n, L), np.float32)
        lib().or_score(self._h, q0, q1, quantile, thre, variant, _p(first), _p(mask), _p(scores))
        return first, mask, scores

    def finetune(self, q0: int, q1: int, mask: np.ndarray, quantile=0.8, thre=0.05, times=0,
                 variant: int = VAR_SEQ):
        n = q1 - q0
        m = np.ascontiguousarray(mask, np.float32)
        final = np.zeros(n, np.uint32)
        rounds = np.zeros(n, np.int32)
        lib().or_finetune(self._h, q0, q1, _p(m), quantile, thre, times, variant, _p(final),
                          _p(rounds))
        return final, rounds

    def finetune_round(self, q: int, top, quantile=0.8, thre=0.05, variant: int = VAR_SEQ):
        """One fine-tune round of query cell `q` on the candidate list `top`:
        returns (new candidate list, per-label scores of the round)."""
        t = np.ascontiguousarray(top, np.uint32)
        out = np.zeros(max(len(t), 1), np.uint32)
        sc = np.full(max(len(t), 1), np.nan, np.float32)
        n = lib().or_finetune_round(self._h, q, _p(t), len(t), quantile, thre, variant, _p(out),
                                    _p(sc))
        return [int(v) for v in out[:n]], sc[:len(t)]


def percentile(arr: np.ndarray, quantile: float) -> float:
    a = np.ascontiguousarray(arr, np.float32)
    return float(lib().or_percentile(_p(a), a.shape[0], quantile))


def minmax(arr: np.ndarray):
    a = np.ascontiguousarray(arr, np.float32)
    mn, mx = _I(), _I()
    lib().or_minmax(_p(a), a.shape[0], C.byref(mn), C.byref(mx))
    return mn.value, mx.value


def rank(codes: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(codes, np.uint16)
    out = np.zeros(c.shape[0], np.float32)
    lib().or_rank(_p(c), c.shape[0], _p(out))
    return out


def corr(q: np.ndarray, r: np.ndarray, variant: int):
    qq = np.ascontiguousarray(q, np.float32)
    rr = np.ascontiguousarray(r, np.float32)
    d = C.c_double()
    f = lib().or_corr(_p(qq), _p(rr), qq.shape[0], variant, C.byref(d))
    return float(f), d.value


def ref_driver_available() -> bool:
    return os.path.exists(REF_DRIVER)


def run_ref_driver(ds, workdir: str, cores: int = 4):
    """Runs the reference's own src/io.cpp (compiled into oracle/_ref) on a
    dataset and returns its preprocessing outputs."""
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "meta.txt"), "w") as f:
        f.write(f"{ds.ref_height} {ds.ref_width} {ds.qry_height} {ds.qry_width} "
                f"{len(ds.codes)} {cores}\n")
    np.ascontiguousarray(ds.ref_data, np.float32).tofile(os.path.join(workdir, "ref_data.f32"))
    np.ascontiguousarray(ds.ref_indices, np.int32).tofile(os.path.join(workdir, "ref_indices.i32"))
    np.ascontiguousarray(ds.ref_indptr, np.int32).tofile(os.path.join(workdir, "ref_indptr.i32"))
    np.ascontiguousarray(ds.qry_data, np.float32).tofile(os.path.join(workdir, "qry_data.f32"))
    np.ascontiguousarray(ds.qry_indices, np.int32).tofile(os.path.join(workdir, "qry_indices.i32"))
    np.ascontiguousarray(ds.qry_indptr, np.int32).tofile(os.path.join(workdir, "qry_indptr.i32"))
    np.ascontiguousarray(ds.celltype, np.int32).tofile(os.path.join(workdir, "celltype.i32"))
    for name, lst in [("ref_genes", ds.ref_genes), ("qry_genes", ds.qry_genes),
                      ("codes", ds.codes), ("cellnames", ds.cellnames)]:
        with open(os.path.join(workdir, name + ".txt"), "w") as f:
            f.write("\n".join(lst) + "\n")
    out = os.path.join(workdir, "out")
    os.makedirs(out, exist_ok=True)
    subprocess.run([REF_DRIVER, workdir, out], check=True, stdout=subprocess.DEVNULL)

    def rd(name, dt):
        return np.fromfile(os.path.join(out, name), dtype=dt)

    res = {}
    for step in (0, 1):
        for part, dt in [("ref_data", np.uint16), ("ref_indices", np.int32),
                         ("ref_indptr", np.int32), ("qry_data", np.uint16),
                         ("qry_indices", np.int32), ("qry_indptr", np.int32)]:
            res[f"step{step}_{part}"] = rd(f"step{step}_{part}.{'u16' if dt == np.uint16 else 'i32'}", dt)
    for name in ["ctidx", "train_idxs", "train_values", "common_genes", "thre_genes", "ctdidx1",
                 "ctdiff1"]:
        res[name] = rd(name + ".u32", np.uint32)
    with open(os.path.join(out, "meta.txt")) as f:
        mu0, mu1, ct = f.read().split()
    res["max_uniq0"], res["max_uniq1"] = int(mu0), int(mu1)
    return res
