#!/usr/bin/env python3
"""Throughput benchmark of the SingleR annotation path on one MI355X.

A step is one pass of the hot path over the whole query batch: the score step
(rank, correlate query cells with reference cells, per-label 0.8 quantile,
first label and candidate mask) followed by the fine-tune loop, with the coded
inputs already resident in HBM.  `value` is annotated query cells per second
over the timed steps.

Default workload: the 200,000-query x 20,000-reference, 50-label atlas
(synthetic, seeded) with 500,000 extra training cells for the learned
proposal.  The chosen form is the exact INT8-limb score path restricted by the
learned top-4 proposal (the network is trained on the GPU beforehand, on the
plain method's final labels of the training cells; training is not timed but
reported).  The plain FP32 path and the exact path without the proposal are
measured beside it.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {"hbm_GBps": 8000.0, "fp32_tflops": 157.3, "fp16_tflops": 2500.0,
               "int8_tops": 5000.0}


def _hip_sync():
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def calibrate_copy(nbytes: int = 256 << 20):
    """One device-to-device copy of `nbytes` (a kernel of known traffic: it
    reads and writes nbytes), used to scale the HBM byte counters."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes))
    hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes))
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def host_cpus():
    """Threads the CPU baseline uses and what the host is: the CPUs this
    process may run on (nproc), capped by OMP_NUM_THREADS where the box sets
    it (its CPU share per GPU), and the CPU model."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cap = os.environ.get("OMP_NUM_THREADS")
    cores = min(nproc, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else nproc
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, nproc, model


def cpu_baseline(ds, n_sample: int, cores: int):
    """The oracle (a CPU restatement of the reference's score + fine-tune steps)
    on the first n_sample query cells."""
    from oracle import oracle as O
    import numpy as np
    orc = O.Oracle(ds, cores)
    t0 = time.perf_counter()
    _, mask, _ = orc.score(0, n_sample, 0.8, 0.05, O.VAR_SEQ)
    orc.finetune(0, n_sample, mask, 0.8, 0.05, 0, O.VAR_SEQ)
    dt = time.perf_counter() - t0
    return {"value": n_sample / dt, "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"first {n_sample} of {ds.qry_height} query cells, score + fine-tune, "
                      f"{dt:.1f} s on {cores} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="learned")
    ap.add_argument("--nq", type=int, default=None, help="override the query cell count")
    ap.add_argument("--form", "--path", dest="form", default="proposal",
                    choices=["plain", "exact_int8", "fp32_mfma", "fp16", "proposal"],
                    help="the chosen form measured as `value` (proposal = exact_int8 restricted "
                         "by the learned proposal); plain is always measured beside it")
    ap.add_argument("--train-cells", type=int, default=None,
                    help="training cells for the proposal network (default: the config's, "
                         "else 500,000)")
    ap.add_argument("--finetune-fuse", type=int, default=None,
                    help="largest fine-tune group (cells) of the fused FP32 form, 0 to 4 "
                         "(default: the engine's, 4)")
    ap.add_argument("--finetune-chain", type=int, default=None,
                    help="longest chain of fine-tune rounds launched without a host read-back "
                         "(0 = none; default: the engine's, 0)")
    ap.add_argument("--finetune-tail", type=int, default=None,
                    help="most active cells of a fine-tune round handed to the persistent tail "
                         "kernel (0 = never; default: the engine's, 64)")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--hidden", type=int, default=512,
                    help="width of the proposal network's two hidden layers (multiples of 256 take the pipelined layer kernel)")
    ap.add_argument("--proposal-k", type=int, default=4)
    ap.add_argument("--min-margin", type=float, default=0.2)
    ap.add_argument("--net-cache", default=None,
                    help="load the proposal network's parameters from this .npy file if it "
                         "exists (skipping training), else save them there after training")
    ap.add_argument("--cpu-sample", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cores", type=int, default=None,
                    help="host threads (default: nproc, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end leg (host query CSR -> host labels)")
    ap.add_argument("--only-chosen", action="store_true",
                    help="measure only the chosen form (for profiler runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: query cells shard across GPUs by design, and "
              f"this benchmark measures one GPU's share (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import dataclasses
    import numpy as np
    from cusingler_amd import engine, synth
    auto_cores, nproc, cpu_model = host_cpus()
    if args.cores is None:
        args.cores = auto_cores
    if args.calibrate_copy:
        calibrate_copy()

    chosen = args.form
    cfg = synth.CONFIGS[args.config]
    cached = chosen == "proposal" and args.net_cache and os.path.exists(args.net_cache)
    if cached:
        cfg = dataclasses.replace(cfg, n_train=0)  # training cells are drawn after the atlas
    elif chosen == "proposal" and (args.train_cells is not None or cfg.n_train == 0):
        cfg = dataclasses.replace(cfg, n_train=args.train_cells or 500000)
    elif chosen != "proposal":
        cfg = dataclasses.replace(cfg, n_train=0)
    t0 = time.perf_counter()
    ds = synth.generate(cfg, n_qry=args.nq)
    t_gen = time.perf_counter() - t0
    print(f"[bench] generated {args.config}: {ds.qry_height} query x {ds.ref_height} reference "
          f"cells (+{cfg.n_train} training cells) in {t_gen:.1f} s", file=sys.stderr, flush=True)

    eng = engine.Engine(0)
    if args.finetune_fuse is not None:
        eng.set_finetune_fuse(args.finetune_fuse)
    if args.finetune_chain is not None:
        eng.set_finetune_chain(args.finetune_chain)
    if args.finetune_tail is not None:
        eng.set_finetune_tail(args.finetune_tail)
    training = None
    if cached:
        net = np.load(args.net_cache, allow_pickle=False)
        training = {"loaded_from": args.net_cache}
    elif chosen == "proposal":
        # label the training cells with the plain method, then train the
        # network on the GPU (not timed)
        t0 = time.perf_counter()
        tprep = engine.Prep(synth.training_set(ds), args.cores)
        eng.copyin_score(tprep.generate(0))
        eng.copyin(tprep.generate(1))
        del tprep
        t1 = time.perf_counter()
        eng.get_label(0.8, 0.05, "plain", outputs=False)
        ylab, _ = eng.finetune(0.8, 0.05, 0)
        _hip_sync()
        t2 = time.perf_counter()
        eng.mlp_init(args.hidden, seed=1)
        losses = eng.mlp_train(ylab.astype(np.int32), epochs=args.epochs, batch=1024, lr=1e-3,
                               seed=2)
        t3 = time.perf_counter()
        training = {"cells": int(ylab.shape[0]), "labels_from": "plain finalLabel",
                    "network": f"G_s -> {args.hidden} -> {args.hidden} -> {len(ds.codes)} (GELU), BF16 MFMA, "
                               "FP32 Adam",
                    "epochs": args.epochs, "batch": 1024, "epoch_loss": [round(float(v), 5) for v in losses],
                    "prep_s": round(t1 - t0, 2), "plain_labelling_s": round(t2 - t1, 3),
                    "train_s": round(t3 - t2, 3)}
        print(f"[bench] proposal network trained on {ylab.shape[0]} cells in {t3 - t2:.2f} s, "
              f"losses {losses}", file=sys.stderr, flush=True)
        net = eng.mlp_params()
        if args.net_cache:
            np.save(args.net_cache, net, allow_pickle=False)
        eng.destroy_score()
        eng.destroy()
        ds.train = None

    prep = engine.Prep(ds, args.cores)
    c0 = prep.generate(0)
    c1 = prep.generate(1)
    print(f"[bench] preprocessing {prep.train_seconds:.1f} s, G_s={c0.width} G_f={c1.width}",
          file=sys.stderr, flush=True)
    # reference training (trainData) on the device, against the host's
    rmap = prep.ref_map(ds.ref_width)
    eng.train(ds, rmap)  # warm
    t0 = time.perf_counter()
    tr = eng.train(ds, rmap)
    dev_train_s = time.perf_counter() - t0
    train_line = {"device_s": round(dev_train_s, 4),
                  "host_prep_s": round(prep.train_seconds, 4), "host_threads": args.cores,
                  "equal_to_host": bool(all(np.array_equal(tr[k], prep.table(i)) for i, k in
                                            enumerate(["train_idxs", "train_values",
                                                       "common_genes", "thre_genes"]))),
                  "note": "device: upload of the reference CSR, medians, pair windows and gene "
                          "sets to the host; host: csg_prep_train (intersection, grouping, "
                          "training, query loading)"}
    eng.copyin_score(c0)
    eng.copyin(c1)  # fine-tune inputs; the candidate mask comes from the device
    Nq, Nr, Gs = ds.qry_height, ds.ref_height, c0.width
    if cached:
        eng.mlp_init(args.hidden, seed=1)
        eng.set_mlp_params(net)

    def measure(form):
        path = "exact_int8" if form == "proposal" else form
        eng.set_proposal(args.proposal_k if form == "proposal" else 0, args.min_margin)

        def step():
            eng.get_label(0.8, 0.05, path, outputs=False)
            return eng.finetune(0.8, 0.05, 0)

        for i in range(args.warmup):
            step()
        # timed steps: HIP events on the contraction phase only (its live launch
        # time feeds the roofline); the other phases' events would cost ~1%
        eng.set_phase_timing(1)
        eng.phase_times(reset=True)
        _hip_sync()
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
            print(f"[bench] {form} step {i + 1}/{args.steps} done", file=sys.stderr, flush=True)
        _hip_sync()
        dt = time.perf_counter() - t0
        ph = eng.phase_times(reset=True)
        # fine-tune rounds of the last timed step that ran in chains (the
        # labelled pass below launches its rounds one by one to time them)
        chained = {f"timed_step_{k}": v for k, v in eng.counters().items()
                   if k in ("ft_chained_rounds", "ft_chain_halts")}
        # labels (and counters) of one untimed pass for the agreement check;
        # the per-phase breakdown is taken from this pass, every phase timed
        # (the device fine-tune's kernel families too: mode 3)
        eng.set_phase_timing(3)
        first, _, _ = eng.get_label(0.8, 0.05, path, outputs=True)
        cnt = eng.counters()
        final, _ = eng.finetune(0.8, 0.05, 0)
        cnt.update({k: v for k, v in eng.counters().items() if k.startswith("ft_")})
        cnt.update(chained)
        ph1 = eng.phase_times(reset=True)
        eng.set_phase_timing(2)
        return dict(value=Nq * args.steps / dt, ms=dt * 1e3 / args.steps, phases=ph, phases1=ph1,
                    counters=cnt, first=first, final=final)

    def measure_e2e(form, steps, warmup, pipelined, narrow=False):
        """Query CSR on the host (float values, pinned) -> first and final labels
        on the host: the upload and device value coding of the query
        (csg_copyin_query_raw, both gene sets), the score step, the fine-tune and
        both label read-backs, per step.  The reference stays loaded (its
        training, coding and upload are one-time work, outside the window).
        `narrow`: the query's column indices held as uint16, as the h5ad
        reader produces them for widths up to 65536 (6 instead of 8 bytes per
        stored value cross PCIe)."""
        path = "exact_int8" if form == "proposal" else form
        eng.set_proposal(args.proposal_k if form == "proposal" else 0, args.min_margin)
        pin = engine.Pinned()
        qd = pin.copy(ds.qry_data)
        qi = pin.copy(np.asarray(ds.qry_indices, np.uint16 if narrow else np.int32))
        qp = pin.copy(np.asarray(ds.qry_indptr, np.int32))
        m0, m1 = prep.query_map(0, ds.qry_width), prep.query_map(1, ds.qry_width)

        def step():
            if pipelined:
                # the next batch (here: the same arrays again) uploads while
                # this one is coded and annotated
                eng.prefetch_query_raw(qd, qi, qp)
            eng.copyin_query_raw(qd, qi, qp, ds.qry_width, m0, m1)
            first, _, _ = eng.get_label(0.8, 0.05, path, outputs="first")
            final, _ = eng.finetune(0.8, 0.05, 0)
            return first, final

        if pipelined:
            eng.prefetch_query_raw(qd, qi, qp)

        for _ in range(warmup):
            step()
        eng.set_phase_timing(0)
        _hip_sync()
        t0 = time.perf_counter()
        for i in range(steps):
            first, final = step()
        _hip_sync()
        dt = time.perf_counter() - t0
        eng.set_phase_timing(2)
        eng.phase_times(reset=True)
        step()
        ph = eng.phase_times(reset=True)
        eng.synchronize()
        if pipelined:
            eng.copyin_query_raw(qd, qi, qp, ds.qry_width, m0, m1)  # consume the last prefetch
        pin.free()
        nnz = int(ds.qry_indptr[-1])
        up_ms = ph["query_coding"][0]
        qbytes = (6 if narrow else 8) * nnz + 4 * (Nq + 1)
        return {"value": round(Nq * steps / dt, 1), "unit": "cells/s", "form": form,
                "ms_per_step": round(dt * 1e3 / steps, 3), "steps": steps,
                "pipelined": pipelined,
                "column_indices": "uint16" if narrow else "int32",
                "query_upload_and_coding_ms": round(up_ms, 3),
                "query_bytes": qbytes,
                "upload_and_coding_GBps": round(qbytes / (up_ms * 1e-3) / 1e9, 1)
                if up_ms > 0 else None,
                "labels_equal_device_resident": bool(
                    np.array_equal(first, res[form]["first"]) and
                    np.array_equal(final, res[form]["final"])),
                "window": "float query CSR in pinned host memory -> first and final labels "
                          "on the host (query H2D + device value coding of both gene sets + "
                          "score step + fine-tune + label D2H); reference resident" +
                          ("; the next batch's H2D overlaps this batch's annotation"
                           if pipelined else "; upload, then annotation")}

    def measure_host_coding(form, steps):
        """The reference's own structure (Pipeline::score then ::finetune,
        src/pipeline.cpp:47-111): host value coding of both steps
        (generateDenseMatrix, on args.cores threads), copyin_score / copyin
        uploads, score step, fine-tune, labels to the host."""
        path = "exact_int8" if form == "proposal" else form
        eng.set_proposal(args.proposal_k if form == "proposal" else 0, args.min_margin)
        t_code = 0.0
        _hip_sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            k0 = prep.generate(0)
            t_code += time.perf_counter() - a
            eng.copyin_score(k0)
            eng.get_label(0.8, 0.05, path, outputs="first")
            a = time.perf_counter()
            k1 = prep.generate(1)
            t_code += time.perf_counter() - a
            eng.copyin(k1)
            eng.finetune(0.8, 0.05, 0)
        _hip_sync()
        dt = time.perf_counter() - t0
        return {"value": round(Nq * steps / dt, 1), "unit": "cells/s", "form": form,
                "ms_per_step": round(dt * 1e3 / steps, 3), "steps": steps,
                "host_coding_ms_per_step": round(t_code * 1e3 / steps, 3),
                "host_threads": args.cores,
                "window": "reference structure: host value coding of reference and query for "
                          "both steps, copyin_score / copyin (reference re-uploaded and "
                          "re-ranked), score step, fine-tune, labels to the host"}

    forms = [chosen] if args.only_chosen else list(dict.fromkeys(
        [chosen, "plain"] + (["exact_int8"] if chosen == "proposal" else [])))
    res = {f: measure(f) for f in forms}
    e2e = None
    if not args.no_e2e:
        print("[bench] end-to-end leg", file=sys.stderr, flush=True)
        e2e = {f: measure_e2e(f, max(1, args.steps // 2), 1, True)
               for f in ([chosen] if args.only_chosen else [chosen, "plain"])}
        e2e["serial_" + chosen] = measure_e2e(chosen, max(1, args.steps // 2), 1, False)
        if ds.qry_width <= 65536:
            e2e["uint16_columns_" + chosen] = measure_e2e(chosen, max(1, args.steps // 2), 1, True,
                                                          narrow=True)
        e2e["host_coding_" + chosen] = measure_host_coding(chosen, 2)
    r = res[chosen]
    steps = args.steps
    corr_ms, corr_n = r["phases"]["corr"]
    if chosen == "proposal":
        # correlations the proposal-restricted score step computed per pass
        corr_per_step = r["counters"]["prop_correlations"]
        ops_desc = "2 * G_s * (correlations of the launch's (cell, label) pairs)"
    else:
        corr_per_step = Nq * Nr
        ops_desc = "2 * G_s * query_rows_in_chunk * reference_cells"
    logical_ops = 2.0 * corr_per_step * Gs * steps          # one logical MAC = 2 ops
    achieved = logical_ops / (corr_ms * 1e-3) / 1e12 if corr_ms > 0 else 0.0
    if chosen in ("exact_int8", "proposal"):
        # folded limbs up to 1,920 genes, three partials beyond
        kname = ("k_gemm_int8w" if Gs <= 1920 else "k_gemm_int8x") + (
            "_group" if chosen == "proposal" else "")
        peak, unit = MI355X_PEAK["int8_tops"], "TOP/s"
    elif chosen == "fp16":
        peak, unit, kname = MI355X_PEAK["fp16_tflops"], "TFLOP/s", "k_gemm_f16"
    elif chosen == "fp32_mfma":
        peak, unit, kname = MI355X_PEAK["fp32_tflops"], "TFLOP/s", "k_gemm_f32mfma"
    else:
        peak, unit, kname = MI355X_PEAK["fp32_tflops"], "TFLOP/s", "k_gemm_plain"
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 5), "traffic": None, "kernel": kname,
                "launches": corr_n, "avg_launch_ms": round(corr_ms / max(corr_n, 1), 4),
                "ops_per_launch": ops_desc}
    if chosen in ("exact_int8", "proposal"):
        # INT8 limb products per logical multiply-add: three in the
        # three-product form (score gene sets of up to 1,008 genes), four in
        # the folded form
        products = 3 if Gs <= 1008 else 4
        roofline["limb_products"] = products
        roofline["limb_mfma_utilisation"] = round(products * achieved / peak, 5)
    workload = (f"{args.config}: {Nq} query x {Nr} reference cells, {len(ds.codes)} labels, "
                f"G_s={Gs}, G_f={c1.width}")
    # HBM bytes per launch from the committed counter runs (FETCH_SIZE and
    # WRITE_SIZE, scaled by a copy kernel of known size), when they cover this
    # workload and kernel
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f)
        for t in (tr if isinstance(tr, list) else [tr]):
            if t.get("kernel") == kname and t.get("workload") == workload:
                roofline["traffic"] = round(t["traffic_bytes_per_launch"])
                roofline["traffic_source"] = t.get("source")
    except (OSError, ValueError, KeyError):
        pass
    # achieved HBM GB/s of the bandwidth-bound score-step stages (algorithmic
    # bytes over the phase time of the labelled pass) and the fine-tune's work
    Kp = (Gs + 63) // 64 * 64
    nnz0 = int(c0.qry_indptr[-1])
    c = r["counters"]
    feat = 2 * Kp if chosen == "proposal" else 0
    limbs = 4 * Kp if chosen in ("exact_int8", "proposal") else 0
    # rank: CSR codes + indices read (6 B / nonzero) and the row pointers,
    # the int16 rank row and its inverse root written, the limb planes (read
    # the ranks, write two int8 planes) and the network features
    rank_bytes = 6 * nnz0 + Nq * (8 + 2 * Kp + 8 + limbs + feat)
    scores_read = c["prop_correlations"] if chosen == "proposal" else Nq * Nr
    select_bytes = 4 * scores_read + 4 * Nq * len(ds.codes)

    def stage(bytes_, ms):
        gbs = bytes_ / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {"bytes": int(bytes_), "ms": round(ms, 4), "GB/s": round(gbs, 1),
                "frac_of_hbm_peak": round(gbs / MI355X_PEAK["hbm_GBps"], 4)}

    stages = {"rank": stage(rank_bytes, r["phases1"]["rank"][0]),
              "select": stage(select_bytes, r["phases1"]["select"][0])}
    pg = c.get("ft_pair_genes", 0)
    ft_ms = r["phases1"]["finetune"][0]
    ft_line = {"pair_genes": pg, "ops": 2 * pg, "bytes": 6 * pg, "ms": round(ft_ms, 3),
               "achieved_TOPs": round(2 * pg / (ft_ms * 1e-3) / 1e12, 3) if ft_ms > 0 else 0.0,
               "families_ms": {k: round(r["phases1"][k][0], 3)
                               for k in ("ft_rank", "ft_corr", "ft_select")},
               "definition": "ops = 2 x sum over rounds of cells x reference rows x |G_sub|; "
                             "bytes = 6 x the same (int16 ranks of both rows + FP32 score)"}
    plain = res.get("plain", r)
    match = {"first": float((r["first"] == plain["first"]).mean()),
             "final": float((r["final"] == plain["final"]).mean())}

    out = {
        "metric": "annotated query cells/sec on 1 MI355X, plain vs chosen form, at label-match %",
        "value": round(r["value"], 1), "unit": "cells/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(r["ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"plain": "f32", "fp32_mfma": "f32", "exact_int8": "int8", "fp16": "f16",
                  "proposal": "int8"}[chosen],
        "data": "synthetic (seeded NB counts, log2(1+x), float32 CSR)",
        "config": {"workload": workload, "form": chosen, "query_cells": Nq, "reference_cells": Nr,
                   "labels": len(ds.codes), "score_genes": Gs, "finetune_genes": c1.width,
                   "network": (f"input: doubled ranks over the score gene set ({Gs} genes), "
                               f"{args.hidden} -> {args.hidden} hidden (GELU), {len(ds.codes)} "
                               f"logits, top-{args.proposal_k}")
                   if chosen == "proposal" else None},
        "plain": {"value": round(plain["value"], 1), "ms_per_step": round(plain["ms"], 3),
                  "phases_ms": {k: round(v[0], 3) for k, v in plain["phases1"].items()}},
        "speedup_vs_plain": round(r["value"] / plain["value"], 3),
        "label_match_vs_plain": match,
        "roofline": roofline,
        "phases_ms_per_step": {k: round(v[0], 3) for k, v in r["phases1"].items()},
        "phases_source": "one untimed pass with every phase timed (the timed steps time the "
                         "contraction phase only)",
        "counters": r["counters"],
        "stages": stages,
        "finetune": ft_line,
        "end_to_end": e2e,
        "reference_training": train_line,
    }
    if "exact_int8" in res and chosen != "exact_int8":
        e = res["exact_int8"]
        out["exact_int8"] = {"value": round(e["value"], 1), "ms_per_step": round(e["ms"], 3),
                             "phases_ms": {k: round(v[0], 3) for k, v in e["phases1"].items()},
                             "label_match_vs_plain": {
                                 "first": float((e["first"] == plain["first"]).mean()),
                                 "final": float((e["final"] == plain["final"]).mean())}}
    if chosen == "proposal":
        c = r["counters"]
        out["proposal"] = dict(training or {}, k=args.proposal_k, min_margin=args.min_margin,
                               scored_pair_fraction=round(c["prop_scored_pairs"] / (Nq * len(ds.codes)), 5),
                               proposal_only_cells=c["prop_cells"],
                               fallback_margin_cells=c["prop_fallback_margin"],
                               fallback_edge_cells=c["prop_fallback_check"],
                               widened_cells=c["prop_widened"])
    out["finetune_autotuner"] = eng.finetune_tuning()
    eng.close()
    if not args.no_cpu:
        n = min(args.cpu_sample, Nq)
        print(f"[bench] CPU baseline on {n} cells, {args.cores} threads", file=sys.stderr,
              flush=True)
        out["cpu_baseline"] = cpu_baseline(ds, n, args.cores)
        out["cpu_baseline"]["nproc"] = nproc
        out["cpu_baseline"]["cpu_model"] = cpu_model
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
