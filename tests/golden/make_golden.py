"""Writes the golden fixtures of tests/golden/.

For a few small seeded datasets it stores the inputs and:
  - the reference's own preprocessing outputs (value-coded CSR over both gene
    sets, marker tables, pair windows), produced by the reference's src/io.cpp
    compiled into oracle/_ref/ref_driver (so this script needs
    /root/reference and `make -C oracle ref`);
  - the oracle's score-step and fine-tune results (first / final labels,
    candidate masks, per-label quantile scores, round counts) in the
    reference's FP32 order.

    python tests/golden/make_golden.py [name ...]
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cusingler_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def datasets():
    return {
        "tiny_shared": synth.tiny(seed=7),
        "tiny_filtered": synth.tiny(n_ref_only=7, n_qry_only=5, seed=11),
        "tiny_dirichlet": synth.tiny(n_ref_per_label=9, n_labels=4, n_genes=400, n_qry=40,
                                     props="dirichlet", seed=5),
        "tiny_nan": synth.with_empty_rows(
            synth.tiny(n_ref_per_label=10, n_labels=3, n_genes=320, n_qry=30, seed=21),
            [3, 17], [5, 11]),
        # every stored value shifted: explicit stored zeros (-1.0) and values
        # of both signs (-1.5), which the medians and codes order around the
        # implicit zeros
        "tiny_zeros": shifted(synth.tiny(n_ref_per_label=8, n_labels=3, n_genes=320, n_qry=24,
                                         seed=13), 1.0),
        "tiny_signed": shifted(synth.tiny(n_ref_per_label=8, n_labels=4, n_genes=360, n_qry=24,
                                          seed=17), 1.5),
    }


def shifted(ds, by):
    import dataclasses
    return dataclasses.replace(ds, ref_data=(ds.ref_data - np.float32(by)).astype(np.float32),
                               qry_data=(ds.qry_data - np.float32(by)).astype(np.float32))


def main():
    if not O.ref_driver_available():
        raise SystemExit("oracle/_ref/ref_driver is missing: run `make -C oracle ref` where "
                         "/root/reference is present")
    only = set(sys.argv[1:])  # dataset names to (re)write; all when none given
    for name, ds in datasets().items():
        if only and name not in only:
            continue
        ref = O.run_ref_driver(ds, f"/tmp/golden_{name}")
        orc = O.Oracle(ds, 4)
        first, mask, scores = orc.score(0, ds.qry_height, 0.8, 0.05, O.VAR_SEQ)
        final, rounds = orc.finetune(0, ds.qry_height, mask, 0.8, 0.05, 0, O.VAR_SEQ)
        out = dict(
            ref_height=ds.ref_height, ref_width=ds.ref_width, qry_height=ds.qry_height,
            qry_width=ds.qry_width, ref_data=ds.ref_data, ref_indices=ds.ref_indices,
            ref_indptr=ds.ref_indptr.astype(np.int32), qry_data=ds.qry_data,
            qry_indices=ds.qry_indices, qry_indptr=ds.qry_indptr.astype(np.int32),
            celltype=ds.celltype.astype(np.int32), codes=np.array(ds.codes),
            cellnames=np.array(ds.cellnames), ref_genes=np.array(ds.ref_genes),
            qry_genes=np.array(ds.qry_genes),
            first=first, mask=mask, scores=scores, final=final, rounds=rounds)
        for k, v in ref.items():
            out["ref_" + k if not k.startswith("step") else k] = np.asarray(v)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print("wrote", name, {k: np.asarray(v).shape for k, v in out.items() if k in ("first", "final")})


if __name__ == "__main__":
    main()
