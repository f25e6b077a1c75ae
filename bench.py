#!/usr/bin/env python3
"""Benchmark: all-pairs Smith-Waterman GCUPS on one MI355X (default), and the
learned-prefilter runs.

Default workload (one step = one pass of the hot path over the whole batch):
  uneven-length set of 20,000 sequences from a 3-level evolution tree
  (100 roots with lognormal lengths, median 300, clipped to [100, 600];
  10 children at 8% mutation; 20 leaves per child at 3%; sub:indel 4:1),
  all 2.0e8 unique pairs through the length-bucketed persistent gfx950
  kernel, with the epsilon-neighbour lists (d <= eps_max) emitted on the
  device.  Sequences are uploaded once before timing (resident in HBM).

GCUPS = sum over the computed unique pairs of la*lb / step seconds / 1e9.

Also reported:
  roofline     -- the SW kernel's integer-op rate (6 ops per cell, measured
                  with HIP events on the engine's stream) against the packed
                  int16 VALU peak of the chip;
  cpu_baseline -- the reference's own score() (compiled from its source into
                  oracle/_ref/) on a bounded random sample of the same
                  sequences, single thread, as the reference's CPU path runs
                  (cpu_baseline_all_cores: the same on all host cores);
  dbscan       -- wall seconds of the drop-in adaptive_clustering_SW CLI on the
                  same 20,000 sequences (FASTA in, split-set files out);
                  amplicon105k runs it on its 105,000 sequences (eps from
                  100 in steps of 100).

Other configurations:
  --config amplicon105k  105,000 amplicon-like sequences, same measurement;
  --config proposer100k  learned proposer: train on 2M exact-labelled pairs, calibrate on
               held-out sequences, then recall / precision / time of the
               prefilter lists against the exact lists on 100,000 held-out
               sequences (and DBSCAN on both);
  --config mix500k  end-to-end DBSCAN on 500,000 heavy-tailed sequences: exact
               lists + DBSCAN against prefilter lists + DBSCAN;
  --config consumers  wall time of the drop-in compareSW, silhouette and
               find_closest_sequence_SW CLIs on two 10,000-sequence sets;
  --config vec  the L1 / L2 vector path: epsilon lists of 200,000 clustered
               5-d projections (L2) and of 1024-d 5-mer count vectors of the
               20,000-sequence tree set (L1), fasta2kmer on that set, and the
               drop-in cluster_dbscan_pca / cluster_dbscan_kmerL1 / fasta2kmer
               CLIs against the compiled reference on bounded samples (outputs
               compared byte for byte there).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config tree20k|amplicon105k|proposer100k|mix500k]
"""
import argparse
import ctypes
import importlib.util
import json
import os
import platform
import random
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "mnhn-tree-tools_amd")
sys.path.insert(0, PKG_DIR)
import synth  # noqa: E402
from proposer_train import (pair_set_recall, same_partition, lists_subset_and_recall,  # noqa: E402
                            proposer100k_workload, mix500k_workload)

WORKLOADS = {
    "tree20k": "uneven-length all-pairs SW: 20,000 seqs, 3-level tree, lognormal lengths "
               "median 300 in [100,600]",
    "amplicon105k": "amplicon-scale all-pairs SW: 105,000 seqs (500 ancestors x 200 at 2-10% "
                    "mutation + 5% noise), lengths uniform in [150,300]",
}

# MI355X: 256 CUs x 64 lanes x 2 (packed int16) x 2.4 GHz
PEAK_INT16_TOPS = 256 * 64 * 2 * 2.4e9 / 1e12
OPS_PER_CELL = 6
# instruction rooflines of the kernel forms, at 64 lanes per CU-clock (packed 16-bit ops issue
# every 4 cycles per wave64 on gfx950, measured): the exact fp16 form retires 3.5 packed DP
# instructions per 2 cells (max3 and the clamp fold two of the 6 algorithmic ops each), the
# packed int16 form 5 per 2 cells (saturating adds fold the floor)
FP16_FORM_ROOFLINE_GCUPS = 256 * 64 * 2.4e9 * 2 / 3.5 / 1e9
INT16_FORM_ROOFLINE_GCUPS = 256 * 64 * 2.4e9 * 2 / 5 / 1e9


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "mnhn_tree_tools_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mnhn_tree_tools_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(seqs, sample_n, seed=7, threads=1, port=False):
    """Reference score() (oracle/_ref/libref_sw.so, compiled from the
    reference's smith-waterman.c with its makefile flags) over all unique
    pairs of a random sample on `threads` host threads (pairs split
    statically); the oracle's clean-room C restatement (same flags) with
    port=True or where the reference build is absent."""
    rng = random.Random(seed)
    sample = rng.sample(range(len(seqs)), sample_n)
    sub = [seqs[i] for i in sample]
    bufs = [ctypes.create_string_buffer(s, len(s) + 1) for s in sub]
    ptrs = (ctypes.c_char_p * len(sub))(*[ctypes.cast(b, ctypes.c_char_p) for b in bufs])
    lens = (ctypes.c_size_t * len(sub))(*[len(s) for s in sub])
    ref = os.path.join(ROOT, "oracle", "_ref", "libref_sw.so")
    if os.path.exists(ref) and not port:
        lib, fn, kind = ctypes.CDLL(ref), "ref_cpu_allpairs", "reference"
    else:
        lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle_sw.so"))
        fn, kind = "oracle_cpu_allpairs", "port"
    f = getattr(lib, fn)
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    cs = ctypes.c_uint64(0)
    secs = f(ptrs, lens, len(sub), threads, 0, None, ctypes.byref(cs))
    L = np.array([len(s) for s in sub], dtype=np.float64)
    cells = 0.5 * (L.sum() ** 2 - (L * L).sum())
    return {"value": cells / secs / 1e9, "unit": "GCUPS", "cores": threads, "kind": kind,
            "sample": f"all {sample_n * (sample_n - 1) // 2} unique pairs of {sample_n} "
                      f"random sequences of the workload ({cells:.3g} cells, {secs:.1f} s); "
                      f"host: {cpu_model()}, nproc {os.cpu_count()}"}


def profiled_traffic(config, strip, key="traffic_bytes_launch"):
    """HBM bytes of the SW kernels from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (calibrated against a plain copy
    kernel): `traffic_bytes_launch` = the roofline's kernel in one step (the
    launch that runs the fp16 and int16 forms' items), `traffic_bytes_per_step`
    = every SW launch of the step; None if there is no summary for it or it
    was taken at another strip width than the step ran (`traffic_strip`)."""
    path = os.path.join(ROOT, "profiles", {"tree20k": "sw_lds_mixed_kernel_tree20k.md"}.get(config, "-"))
    vals = {}
    try:
        with open(path) as f:
            for line in f:
                k, sep, v = line.partition(":")
                if sep and k in (key, "traffic_strip"):
                    vals[k] = float(v)
    except OSError:
        pass
    if vals.get("traffic_strip") != strip:
        return None
    return vals.get(key)


def run_dbscan(seqs, eps_start, eps_inc, minpts):
    exe = os.path.join(PKG_DIR, "bin", "adaptive_clustering_SW")
    with tempfile.TemporaryDirectory() as tmp:
        fasta = os.path.join(tmp, "in.fasta")
        synth.write_fasta(fasta, seqs)
        t0 = time.perf_counter()
        r = subprocess.run([exe, fasta, str(eps_start), str(eps_inc), str(minpts),
                            os.path.join(tmp, "layer"), "1"], capture_output=True, timeout=900)
        secs = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(r.stdout.decode()[-2000:] + r.stderr.decode()[-2000:])
        layers = sum(1 for f in os.listdir(tmp) if f.startswith("layer"))
    return {"exact_s_per_run": secs, "eps_start": eps_start, "eps_inc": eps_inc,
            "minpts": minpts, "layers_written": layers,
            "stdout_tail": r.stdout.decode().splitlines()[-1:]}


# ------------------------------------------------------ learned proposer runs

def proposer_eval(args, pkg, log, config):
    """Train the encoder, then time exact and learned-prefilter epsilon lists
    + DBSCAN on the evaluation set of `config` (proposer100k or mix500k).
    Returns (results, workload text, pairwise-kernel roofline, sequences)."""
    eps = args.eps if args.eps else (300 if config == "proposer100k" else 200)
    minpts = 5
    if config == "proposer100k":
        enc, info, seqs, workload = proposer100k_workload(
            pkg, log, eps=eps, pairs=args.pairs, epochs=args.epochs, train_anc=args.train_anc,
            train_per=args.train_per, use_cosine=not args.no_cosine, recall=args.recall_target or 0.9999,
            exact_below=max(args.exact_below, 0), n=args.n or 100_000)
    else:
        enc, info, seqs, workload = mix500k_workload(
            pkg, log, eps=eps, pairs=args.pairs, epochs=args.epochs, train_draws=args.train_draws,
            train_n=args.train_n, use_cosine=not args.no_cosine, recall=args.recall_target or 0.999999,
            exact_below=args.exact_below if args.exact_below >= 0 else 100, n=args.n or 500_000)
    n = len(seqs)
    L = np.array([len(x) for x in seqs], dtype=np.float64)
    log(f"evaluation set: n={n}, mean length {L.mean():.0f}, max {int(L.max())}; eps {eps}")
    hist_edges = [0, 100, 150, 200, 300, 400, 600, 800, 1000, 1500, 1 << 30]
    length_hist = {f"{a}-{b - 1}" if b < (1 << 30) else f">={a}": int(((L >= a) & (L < b)).sum())
                   for a, b in zip(hist_edges[:-1], hist_edges[1:])}
    # each path runs on a context of its own (sequences uploaded, tuning read
    # from the cache), so neither finds device buffers the other sized
    eng = pkg.Engine(seqs)
    if not args.no_autotune:
        log(f"autotune: {eng.autotune(use_cache=True)}")
    # learned prefilter + exact verification
    peng = pkg.Engine(seqs)
    if not args.no_autotune:
        peng.autotune(use_cache=True)
    t0 = time.perf_counter()
    prp, pcol, pdist, pst = enc.prefilter_csr(peng, eps)
    pref_lists_s = time.perf_counter() - t0
    vt = peng.last_timing()
    peng.close()
    verify_forms = {name: {"cells": vt["form_cells"][f], "kernel_ms": vt["form_ms"][f],
                           "gcups": vt["form_cells"][f] / max(vt["form_ms"][f], 1e-9) / 1e6}
                    for f, name in enumerate(("fp16", "int16", "int32")) if vt["form_cells"][f] > 0}
    log(f"prefilter lists: {pref_lists_s:.2f} s {pst}, nnz {len(pcol)}; verification by form {verify_forms}")
    # exact path: all-pairs epsilon lists (length-bound pruned) + DBSCAN
    t0 = time.perf_counter()
    erp, ecol, edist = eng.eps_csr(eps)
    exact_lists_s = time.perf_counter() - t0
    et = eng.last_timing()
    log(f"exact eps lists: {exact_lists_s:.2f} s (kernel {et['kernel_ms'] / 1e3:.2f} s, "
        f"{et['pairs']:.3g} pairs, {et['cells']:.3g} cells), nnz {len(ecol)}")
    unpruned = None
    if args.exact_unpruned:
        # the same lists from every unique pair (no length-bound pruning)
        os.environ["MNHN_SW_LENGTH_PRUNE"] = "0"
        try:
            t0 = time.perf_counter()
            urp, ucol, udist = eng.eps_csr(eps)
            unpruned = {"lists_s": time.perf_counter() - t0,
                        "kernel_s": eng.last_timing()["kernel_ms"] / 1e3,
                        "pairs": eng.last_timing()["pairs"],
                        "identical_lists": bool(np.array_equal(urp, erp) and np.array_equal(ucol, ecol)
                                                and np.array_equal(udist, edist))}
            del urp, ucol, udist
        finally:
            del os.environ["MNHN_SW_LENGTH_PRUNE"]
        log(f"exact all-pairs lists (no length pruning): {unpruned}")
    _, _, missed = pair_set_recall(prp, pcol, pdist, erp, ecol, edist)
    subset, recall = lists_subset_and_recall(prp, pcol, pdist, erp, ecol, edist)
    precision = pst["neighbours"] / max(pst["candidates"], 1)
    log(f"recall {recall:.6f} (proposed subset of exact: {subset}), precision {precision:.4f}")
    if missed:
        mi = np.array(missed)
        lm = np.maximum(L[mi[:, 0]], L[mi[:, 1]]).astype(np.int64)
        log(f"missed pairs in sampled rows: {len(mi)}; longer length quantiles "
            f"{np.quantile(lm, [0, 0.25, 0.5, 0.75, 1]).tolist()}; first "
            f"{[(int(a), int(b), int(L[a]), int(L[b])) for a, b in mi[:8]]}")
    all_pairs = n * (n - 1) / 2
    # the pairwise kernel's executed MFMA work: every band tile runs the
    # 128 x 128 dot products of the 128-d embeddings (bf16), the tiles with a
    # stage-1 survivor (counted on the device) also those of the 1024-d counts
    # (int8 MFMA, or bf16 when a count exceeds 127).  Peak: the time both
    # parts take at their dense peaks (bf16 2.5 PFLOP/s, int8 5.0 POP/s).
    f1 = 2 * 128 * 128 * 128 * pst["tiles"]
    f2 = 2 * 128 * 128 * 1024 * pst["stage2_tiles"]
    t_peak = f1 / 2.5e15 + f2 / (5.0e15 if pst["counts_int8"] else 2.5e15)
    kernel_s = pst["propose_kernel_ms"] / 1e3
    mfma = {"bound": "mfma", "kernel": "pair_candidates_kernel",
            "achieved": (f1 + f2) / kernel_s / 1e12, "peak": (f1 + f2) / t_peak / 1e12,
            "unit": "TFLOP/s (dense MFMA: 2*128*128*128 bf16 flop per band tile + 2*128*128*1024 "
                    + ("int8" if pst["counts_int8"] else "bf16") + " ops per tile that runs the cosine "
                    "stage; peak = both parts at their published dense peaks)",
            "traffic": None, "band_tiles": pst["tiles"], "cosine_stage_tiles": pst["stage2_tiles"],
            "counts_int8": bool(pst["counts_int8"]), "kernel_ms": pst["propose_kernel_ms"]}
    mfma["frac"] = t_peak / kernel_s
    res = {"eps": eps, "recall": recall, "proposed_subset_of_exact": bool(subset),
           "precision": precision, "candidates": pst["candidates"],
           "neighbour_pairs": pst["neighbours"], "exact_neighbour_pairs": len(ecol) // 2,
           "exact_lists_s": exact_lists_s, "prefilter_lists_s": pref_lists_s,
           "exact_kernel_s": et["kernel_ms"] / 1e3, "exact_pairs_computed": et["pairs"],
           "exact_cells_computed": et["cells"], "all_pairs": all_pairs,
           "all_pairs_cells": 0.5 * (L.sum() ** 2 - (L * L).sum()),
           "verify_gcups": pst["cells"] / max(pst["verify_kernel_ms"], 1e-9) / 1e6,
           "verify_forms": verify_forms,
           "prefilter_phases_ms": pst, "training": info}
    t0 = time.perf_counter()
    ecl = pkg.dbscan_csr(erp, ecol, edist, eps, eps, minpts)
    exact_db_s = time.perf_counter() - t0
    del ecol, edist
    t0 = time.perf_counter()
    pcl = pkg.dbscan_csr(prp, pcol, pdist, eps, eps, minpts)
    pref_db_s = time.perf_counter() - t0
    exact_s = exact_lists_s + exact_db_s
    pref_s = pref_lists_s + pref_db_s
    same = same_partition(ecl, pcl)
    log(f"DBSCAN(eps {eps}, minpts {minpts}): exact {exact_s:.2f} s ({len(ecl)} clusters), "
        f"prefilter {pref_s:.2f} s ({len(pcl)} clusters); same partition: {same}, "
        f"same member order: {ecl == pcl}")
    if unpruned:
        unpruned["dbscan_s"] = unpruned["lists_s"] + exact_db_s
        unpruned["speedup_of_prefilter"] = unpruned["dbscan_s"] / pref_s
        res["exact_all_pairs"] = unpruned
    else:
        # exact all-pairs DBSCAN (every unique pair through the SW kernel, the
        # baseline the workload names) estimated from the length-pruned exact
        # run: all pairs' cells at its measured kernel rate, plus its list and
        # host DBSCAN time
        rate = et["cells"] / max(et["kernel_ms"] / 1e3, 1e-9)
        est = res["all_pairs_cells"] / rate + (exact_lists_s - et["kernel_ms"] / 1e3) + exact_db_s
        res["exact_all_pairs_estimate"] = {"dbscan_s": est, "speedup_of_prefilter": est / pref_s,
                                           "basis": "all unique pairs' cells at the exact path's "
                                                    "measured kernel GCUPS + its list and DBSCAN host time"}
    res.update({"minpts": minpts, "exact_dbscan_s": exact_s, "prefilter_dbscan_s": pref_s,
                "speedup": exact_s / pref_s, "host_dbscan_s": [exact_db_s, pref_db_s],
                "clusters_exact": len(ecl), "clusters_prefilter": len(pcl),
                "identical_clusters": bool(same), "identical_member_order": ecl == pcl,
                "clustered_points": int(sum(len(c) for c in ecl)), "length_histogram": length_hist})
    eng.close()
    enc.close()
    return res, workload, mfma, seqs


def run_proposer_bench(args, pkg, log):
    res, workload, mfma, seqs = proposer_eval(args, pkg, log, args.config)
    n = len(seqs)
    out = {}
    common = {"n_gpus": 1, "steps": 1, "warmup": 0, "scaling": "weak", "vs_baseline": None,
              "data": "synthetic", "dtype": "bf16 (encoder) / f16 exact integers + int16 (verify)",
              "config": {"workload": workload, "n_seqs": n, **res}, "roofline": mfma}
    if args.config == "proposer100k":
        out.update({"metric": "learned-prefilter eps-neighbour recall (held-out)",
                    "value": res["recall"], "unit": "fraction", "higher_is_better": True,
                    "ms_per_step": res["prefilter_lists_s"] * 1e3, **common})
    else:
        out.update({"metric": "DBSCAN s/run exact vs learned-prefilter", "value": res["prefilter_dbscan_s"],
                    "unit": "s", "higher_is_better": False, "ms_per_step": res["prefilter_dbscan_s"] * 1e3,
                    **common})
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(seqs, args.cpu_sample)
    print(json.dumps(out), flush=True)


# ------------------------------------------------- other score() consumers

def run_consumers_bench(args, log):
    """Wall time of the drop-in compareSW, silhouette and
    find_closest_sequence_SW CLIs on two uneven-length sets (two draws of the
    20,000-sequence tree generator, n of each), against the reference's
    score() rate on the same data (its CPU path computes the same cells)."""
    n = args.n or 10000
    a = synth.evolution_tree(seed=2)[::2][:n]
    b = synth.evolution_tree(seed=3)[::2][:n]
    La = np.array([len(x) for x in a], np.float64)
    Lb = np.array([len(x) for x in b], np.float64)
    cross = float(La.sum() * Lb.sum())
    within = float(La.sum() ** 2)
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        fa, fb, fq = (os.path.join(tmp, x) for x in ("a.fasta", "b.fasta", "q.fasta"))
        synth.write_fasta(fa, a)
        synth.write_fasta(fb, b)
        synth.write_fasta(fq, b[:1])
        runs = [("compareSW", [fa, fb, "1"], 2 * cross),  # mean pass + sigma pass
                ("silhouette", [fa, fb, "1"], within + cross),
                ("find_closest_sequence_SW", [fa, fq, "1", "1", os.path.join(tmp, "c.fasta")],
                 float(La.sum() * Lb[0]))]
        for prog, argv, cells in runs:
            exe = os.path.join(PKG_DIR, "bin", prog)
            t0 = time.perf_counter()
            r = subprocess.run([exe] + argv, capture_output=True, timeout=900)
            secs = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(r.stdout.decode()[-500:] + r.stderr.decode()[-500:])
            res[prog] = {"s": secs, "cells": cells, "gcups_wall": cells / secs / 1e9,
                         "stdout": r.stdout.decode().strip()[:120]}
            log(f"{prog}: {secs:.2f} s, {cells / secs / 1e9:.0f} GCUPS (wall), {res[prog]['stdout']}")
    out = {"metric": "SW consumer CLIs wall seconds (compareSW)", "value": res["compareSW"]["s"],
           "unit": "s", "higher_is_better": False, "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": res["compareSW"]["s"] * 1e3, "scaling": "weak", "vs_baseline": None,
           "dtype": "f16 (exact integers) + int16", "data": "synthetic",
           "config": {"workload": f"compareSW / silhouette / find_closest_sequence_SW on two sets of "
                                  f"{n} uneven-length sequences (tree generator draws)",
                      "runs": res}}
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cb = cpu_baseline(a + b, args.cpu_sample)
        out["cpu_baseline"] = cb
        out["config"]["reference_cpu_estimate_s"] = {
            k: v["cells"] / (cb["value"] * 1e9) for k, v in res.items()}
    print(json.dumps(out), flush=True)


# MI355X published fp32 vector peak (an FMA counts 2 flops)
PEAK_FP32_TFLOPS = 157.3
# 1,024 SIMDs x 2.4 GHz x 64 lanes / 6 cycles per pair-dimension (sub 2 + fma 4)
L2_INSTR_PAIR_DIMS_S = 1024 * 2.4e9 * 64 / 6


def _run_cli(argv, cwd, timeout=1800):
    t0 = time.perf_counter()
    r = subprocess.run(argv, cwd=cwd, capture_output=True, timeout=timeout)
    secs = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(" ".join(argv[:1]) + ": " + r.stdout.decode()[-300:] + r.stderr.decode()[-300:])
    return secs, r.stdout


def _outputs(d, prefix):
    return {f: open(os.path.join(d, f), "rb").read() for f in sorted(os.listdir(d)) if f.startswith(prefix)}


def run_vec_bench(args, pkg, log):
    """Epsilon-list kernel rates of the vector engine and the drop-in vector
    CLIs against the compiled reference (oracle/_ref) on bounded samples."""
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    have_ref = os.path.exists(os.path.join(ref_dir, "cluster_dbscan_pca"))
    rng = np.random.default_rng(11)
    # clustered 5-d projections: 2,000 centres with lognormal weights
    n, dims = (args.n or 200_000), 5
    centres = rng.normal(scale=12.0, size=(2000, dims))
    w = rng.lognormal(0.0, 1.0, 2000)
    lab = rng.choice(2000, n, p=w / w.sum())
    pts = (centres[lab] + rng.normal(scale=0.3, size=(n, dims))).astype(np.float32)
    eps_pca, minpts = 0.5, 5
    thr = pkg.vec_threshold(2, eps_pca)
    res = {}
    with pkg.VecEngine(pts) as eng:
        for _ in range(args.warmup):
            eng.eps_csr(2, thr)
        times, kms = [], []
        for k in range(args.steps):
            t0 = time.perf_counter()
            rp, col, val = eng.eps_csr(2, thr)
            times.append(time.perf_counter() - t0)
            kms.append(eng.last_timing()["kernel_ms"])
            log(f"L2 step {k}: {times[-1] * 1e3:.1f} ms (tile kernel {kms[-1]:.1f} ms), nnz {len(col)}")
        nnz = len(col)
        del rp, col, val
    pairs = n * (n + 1) / 2
    kernel_s = float(np.mean(kms)) / 1e3
    res["l2_pca"] = {"n": n, "dims": dims, "eps": eps_pca, "list_entries": nnz,
                     "mean_neighbours": nnz / n, "step_ms": float(np.mean(times)) * 1e3,
                     "kernel_ms": kernel_s * 1e3, "pair_dims_per_s": pairs * dims / kernel_s,
                     "pairs_per_s_step": pairs / float(np.mean(times))}
    # 5-mer count vectors of the tree set (counted on the GPU)
    seqs = synth.evolution_tree()
    t0 = time.perf_counter()
    km = pkg.kmer_frequencies(seqs, 5, False)
    kmer_s = time.perf_counter() - t0
    kv = km.astype(np.float32)
    eps_l1 = 150.0
    with pkg.VecEngine(kv) as eng:
        eng.eps_csr(1, eps_l1)
        t0 = time.perf_counter()
        rp, col, val = eng.eps_csr(1, eps_l1)
        l1_s = time.perf_counter() - t0
        l1_kernel_s = eng.last_timing()["kernel_ms"] / 1e3
    m = len(seqs)
    res["l1_kmer"] = {"n": m, "dims": 1024, "eps": eps_l1, "list_entries": int(len(col)),
                      "step_ms": l1_s * 1e3, "kernel_ms": l1_kernel_s * 1e3,
                      "pair_dims_per_s": m * (m + 1) / 2 * 1024 / l1_kernel_s,
                      "kmer_counts_s": kmer_s}
    log(f"L1 k-mer lists: {l1_s * 1e3:.1f} ms (tile kernel {l1_kernel_s * 1e3:.1f} ms), nnz {len(col)}; "
        f"5-mer table of {m} sequences {kmer_s * 1e3:.1f} ms")
    # flops per pair-dimension: sub + fma (L2) = 3, sub + add (L1) = 2
    achieved = 3 * res["l2_pca"]["pair_dims_per_s"] / 1e12
    res["l1_kmer"]["tflops"] = 2 * res["l1_kmer"]["pair_dims_per_s"] / 1e12
    res["l1_kmer"]["frac_of_published_fp32"] = res["l1_kmer"]["tflops"] / PEAK_FP32_TFLOPS
    out = {"metric": "L2 epsilon-list pair-dimensions per second (vector DBSCAN scans)",
           "value": res["l2_pca"]["pair_dims_per_s"], "unit": "pair-dims/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["l2_pca"]["step_ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"L2 epsilon lists of {n} clustered 5-d projections (2,000 centres), "
                                  f"eps {eps_pca}; L1 lists of 1024-d 5-mer vectors of 20,000 tree sequences",
                      "runs": res},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s (fp32 VALU; per pair-dimension a sub and an fma = 3 flop)",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                        "kernel": "vec_tiles_kernel<2>",
                        # issue-rate bound of the kernel's instruction mix (measured
                        # on gfx950 with an issue-rate probe): per pair-dimension a
                        # v_sub_f32 (2 cycles per wave64) and a v_fmac_f32 (4)
                        "instruction_roofline_pair_dims_s": L2_INSTR_PAIR_DIMS_S,
                        "instruction_frac": res["l2_pca"]["pair_dims_per_s"] / L2_INSTR_PAIR_DIMS_S}}
    # drop-in CLIs end to end, and the reference on samples (outputs compared)
    with tempfile.TemporaryDirectory() as tmp:
        fa = os.path.join(tmp, "p.fasta")
        fp = os.path.join(tmp, "p.txt")
        synth.write_fasta(fa, [b"ACGT"] * n)
        np.savetxt(fp, pts, fmt="%.7g")
        os.makedirs(os.path.join(tmp, "g"))
        secs, so = _run_cli([os.path.join(PKG_DIR, "bin", "cluster_dbscan_pca"), fa, fp, "5", str(eps_pca),
                             str(minpts), "o"], os.path.join(tmp, "g"))
        res["cli_pca"] = {"n": n, "s": secs, "stdout": so.decode().strip()}
        log(f"cluster_dbscan_pca {n}: {secs:.2f} s, {so.decode().strip()}")
        fk = os.path.join(tmp, "k.fasta")
        synth.write_fasta(fk, seqs)
        secs, ktext = _run_cli([os.path.join(PKG_DIR, "bin", "fasta2kmer"), fk, "5", "1", "0"], tmp)
        res["cli_fasta2kmer"] = {"n": m, "s": secs, "bytes": len(ktext)}
        log(f"fasta2kmer {m} seqs k=5: {secs:.2f} s ({len(ktext) / 1e6:.0f} MB)")
        kf = os.path.join(tmp, "k.kmer")
        with open(kf, "wb") as f:
            f.write(ktext)
        os.makedirs(os.path.join(tmp, "gk"))
        secs, so = _run_cli([os.path.join(PKG_DIR, "bin", "cluster_dbscan_kmerL1"), fk, kf, str(eps_l1), "5",
                             "o"], os.path.join(tmp, "gk"))
        res["cli_kmerL1"] = {"n": m, "s": secs, "stdout": so.decode().strip()}
        log(f"cluster_dbscan_kmerL1 {m}: {secs:.2f} s, {so.decode().strip()}")
        if have_ref and not args.no_cpu_baseline:
            cmp = {}
            # L2 scans on a 40,000-point sample: the reference's pass costs
            # sample^2 * dims pair-dimensions (one region query per point)
            ns = min(n, 40000)
            fa2, fp2 = os.path.join(tmp, "s.fasta"), os.path.join(tmp, "s.txt")
            synth.write_fasta(fa2, [b"ACGT"] * ns)
            np.savetxt(fp2, pts[:ns], fmt="%.7g")
            for tag, exe_dir in (("ours", os.path.join(PKG_DIR, "bin")), ("ref", ref_dir)):
                d = os.path.join(tmp, "s_" + tag)
                os.makedirs(d)
                cmp[tag] = _run_cli([os.path.join(exe_dir, "cluster_dbscan_pca"), fa2, fp2, "5", str(eps_pca),
                                     str(minpts), "o"], d)
                cmp[tag + "_files"] = _outputs(d, "o")
            same_pca = cmp["ours"][1] == cmp["ref"][1] and cmp["ours_files"] == cmp["ref_files"]
            ref_rate = ns * ns * dims / cmp["ref"][0]
            log(f"reference cluster_dbscan_pca {ns}: {cmp['ref'][0]:.2f} s; ours {cmp['ours'][0]:.2f} s; "
                f"identical outputs: {same_pca}")
            # fasta2kmer on a 1,000-sequence sample
            fk2 = os.path.join(tmp, "ks.fasta")
            synth.write_fasta(fk2, seqs[:1000])
            rs, rtext = _run_cli([os.path.join(ref_dir, "fasta2kmer"), fk2, "5", "1", "0"], tmp)
            os_, otext = _run_cli([os.path.join(PKG_DIR, "bin", "fasta2kmer"), fk2, "5", "1", "0"], tmp)
            log(f"reference fasta2kmer 1000 seqs: {rs:.2f} s; ours {os_:.2f} s; identical: {rtext == otext}")
            # L1 k-mer DBSCAN on a 4,000-sequence sample
            ks = 4000
            fk3, kf3 = os.path.join(tmp, "k3.fasta"), os.path.join(tmp, "k3.kmer")
            synth.write_fasta(fk3, seqs[:ks])
            with open(kf3, "wb") as f:
                f.write(b"\n".join(ktext.split(b"\n", ks)[:ks]) + b"\n")
            for tag, exe_dir in (("kours", os.path.join(PKG_DIR, "bin")), ("kref", ref_dir)):
                d = os.path.join(tmp, "s_" + tag)
                os.makedirs(d)
                cmp[tag] = _run_cli([os.path.join(exe_dir, "cluster_dbscan_kmerL1"), fk3, kf3, str(eps_l1), "5",
                                     "o"], d)
                cmp[tag + "_files"] = _outputs(d, "o")
            same_k = cmp["kours"][1] == cmp["kref"][1] and cmp["kours_files"] == cmp["kref_files"]
            log(f"reference cluster_dbscan_kmerL1 {ks}: {cmp['kref'][0]:.2f} s; ours {cmp['kours'][0]:.2f} s; "
                f"identical outputs: {same_k}")
            out["cpu_baseline"] = {
                "value": ref_rate, "unit": "pair-dims/s", "cores": 1, "kind": "reference",
                "sample": f"the reference's cluster_dbscan_pca (oracle/_ref, one thread, AVX2+FMA build) on the "
                          f"first {ns} points ({cmp['ref'][0]:.2f} s wall incl. file parsing; one region "
                          f"query per point = {ns}^2 x {dims} pair-dims); host: {cpu_model()}"}
            res["reference_samples"] = {
                "cluster_dbscan_pca": {"n": ns, "ref_s": cmp["ref"][0], "ours_s": cmp["ours"][0],
                                       "identical": same_pca},
                "fasta2kmer": {"n": 1000, "ref_s": rs, "ours_s": os_, "identical": rtext == otext},
                "cluster_dbscan_kmerL1": {"n": ks, "ref_s": cmp["kref"][0], "ours_s": cmp["kours"][0],
                                          "identical": same_k},
                "reference_estimate_full_s": {
                    "cluster_dbscan_pca": n * n * dims / ref_rate,
                    "fasta2kmer": rs * m / 1000,
                    "cluster_dbscan_kmerL1": cmp["kref"][0] * (m / ks) ** 2}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps-max", type=int, default=2500)
    ap.add_argument("--n", type=int, default=0, help="use only the first n sequences")
    ap.add_argument("--cpu-sample", type=int, default=260,
                    help="sequences in the CPU-baseline sample (all their unique pairs; about 10 s single-thread)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-dbscan", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-prefilter", action="store_true",
                    help="tree20k/amplicon105k: skip the learned-prefilter vs exact DBSCAN block (proposer100k)")
    ap.add_argument("--config", default="tree20k",
                    choices=["tree20k", "amplicon105k", "proposer100k", "mix500k", "consumers", "vec"],
                    help="tree20k: 20,000 uneven-length (headline); amplicon105k: 105,000 amplicon-scale; "
                         "proposer100k: learned-proposer training + held-out recall; mix500k: end-to-end "
                         "learned-prefilter DBSCAN vs the exact path")
    ap.add_argument("--eps", type=int, default=0,
                    help="proposer100k/mix500k epsilon (default 300 for proposer100k, 200 for mix500k)")
    ap.add_argument("--pairs", type=int, default=2_000_000, help="proposer100k/mix500k labelled training pairs")
    ap.add_argument("--epochs", type=int, default=3, help="proposer100k/mix500k training epochs")
    ap.add_argument("--train-n", type=int, default=100_000, help="mix500k sequences per training draw")
    ap.add_argument("--train-draws", type=int, default=4, help="mix500k training draws")
    ap.add_argument("--train-anc", type=int, default=20000, help="proposer100k training ancestors")
    ap.add_argument("--train-per", type=int, default=20, help="proposer100k training descendants per ancestor")
    ap.add_argument("--no-cosine", action="store_true", help="proposer100k/mix500k: embedding radius test only")
    ap.add_argument("--exact-unpruned", action="store_true",
                    help="proposer100k/mix500k: also time the exact lists over every unique pair (no length pruning)")
    ap.add_argument("--recall-target", type=float, default=0,
                    help="proposer100k/mix500k: calibrated per-bucket recall of the proposal tests "
                         "(default 0.9999 for proposer100k, 0.999999 for mix500k)")
    ap.add_argument("--exact-below", type=int, default=-1,
                    help="proposer100k/mix500k: pairs whose longer sequence is at most this long skip the "
                         "proposal tests (default 0 for proposer100k, 100 for mix500k)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (torch) 3 times, "
                         "to calibrate FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    # the engine's phase and autotune lines (stderr) keep a long run visibly alive
    os.environ.setdefault("MNHN_TIMING", "1")
    pkg = load_pkg()
    t_start = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - t_start:7.1f} s] {msg}", file=sys.stderr, flush=True)
    if args.config in ("proposer100k", "mix500k"):
        run_proposer_bench(args, pkg, log)
        return
    if args.config == "consumers":
        run_consumers_bench(args, log)
        return
    if args.config == "vec":
        run_vec_bench(args, pkg, log)
        return
    log("generating workload ...")
    seqs = synth.evolution_tree() if args.config == "tree20k" else synth.amplicon_families()
    if args.n:
        seqs = seqs[:args.n]
    L = np.array([len(s) for s in seqs], dtype=np.float64)
    cells = 0.5 * (L.sum() ** 2 - (L * L).sum())
    pairs = len(seqs) * (len(seqs) - 1) // 2
    log(f"n={len(seqs)} pairs={pairs} cells={cells:.3e}; uploading")

    eng = pkg.Engine(seqs)
    tuning = None
    if not args.no_autotune:
        log("autotune (one line per candidate) ...")
        tuning = eng.autotune(use_cache=True)
        log(f"autotune: {tuning}")
    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        eng.eps_count(args.eps_max)
    times, kms = [], []
    form_cells, form_ms = np.zeros(3), np.zeros(3)
    nnz = 0
    for k in range(args.steps):
        t0 = time.perf_counter()
        nnz = eng.eps_count(args.eps_max)
        times.append(time.perf_counter() - t0)
        t = eng.last_timing()
        kms.append(t["kernel_ms"])
        form_cells += np.array(t["form_cells"]) / args.steps
        form_ms += np.array(t["form_ms"]) / args.steps
        log(f"step {k}: {times[-1] * 1e3:.1f} ms (kernels {kms[-1]:.1f} ms: fp16 {t['form_ms'][0]:.1f}, "
            f"int16 {t['form_ms'][1]:.1f}, int32 {t['form_ms'][2]:.1f}), "
            f"{t['cells'] / times[-1] / 1e9:.1f} GCUPS, nnz {nnz}")
    t = eng.last_timing()
    # GCUPS counts the cells of the pairs the step computed: every unique pair,
    # unless eps_max is small enough for the length bound to leave some out
    comp_cells, comp_pairs = t["cells"], t["pairs"]
    # the same step with the int16 form of the LDS kernel only (integer VALU path)
    eng.set_tuning(2, 32, 16)
    t0 = time.perf_counter()
    eng.eps_count(args.eps_max)
    int16_s = time.perf_counter() - t0
    int16_kernel_ms = eng.last_timing()["kernel_ms"]
    log(f"int16-only step: {int16_s * 1e3:.1f} ms, {comp_cells / int16_s / 1e9:.1f} GCUPS")
    eng.close()
    step_s = float(np.mean(times))
    kernel_s = float(np.mean(kms)) / 1e3
    gcups = comp_cells / step_s / 1e9
    peaks = [FP16_FORM_ROOFLINE_GCUPS, INT16_FORM_ROOFLINE_GCUPS, None]
    forms = {}
    mixed = bool(t.get("mixed_launch"))
    if mixed:
        # one launch ran the fp16 form's items and the int16 form's (its time is
        # form_ms[0]): its roofline is the time both forms' cells take at their
        # own instruction rooflines
        c0, c1 = form_cells[0], form_cells[1]
        at_roof = c0 / (peaks[0] * 1e9) + c1 / (peaks[1] * 1e9)
        ms = form_ms[0]
        forms["fp16+int16 (one launch)"] = {
            "cells": c0 + c1, "fp16_cells": c0, "int16_cells": c1, "kernel_ms": ms,
            "gcups": (c0 + c1) / (ms / 1e3) / 1e9,
            "instruction_roofline_gcups": (c0 + c1) / at_roof / 1e9, "frac": at_roof / (ms / 1e3)}
        if form_cells[2] > 0:
            forms["int32"] = {"cells": form_cells[2], "kernel_ms": form_ms[2],
                              "gcups": form_cells[2] / (form_ms[2] / 1e3) / 1e9,
                              "instruction_roofline_gcups": None, "frac": None}
    else:
        for f, name in enumerate(("fp16", "int16", "int32")):
            if form_cells[f] > 0:
                g = form_cells[f] / (form_ms[f] / 1e3) / 1e9
                forms[name] = {"cells": form_cells[f], "kernel_ms": form_ms[f], "gcups": g,
                               "instruction_roofline_gcups": peaks[f],
                               "frac": g / peaks[f] if peaks[f] else None}
    # dominant kernel: the one launch of both forms (sw_lds_mixed_kernel), or
    # the exact fp16 form's launches
    f16 = forms.get("fp16+int16 (one launch)") or forms.get("fp16") or forms.get("int16")
    f16_peak = f16["instruction_roofline_gcups"]
    # the step against each form's own roofline: time at roofline / kernel time
    at_roof_s = sum(forms[k]["cells"] / (forms[k]["instruction_roofline_gcups"] * 1e9)
                    for k in forms if forms[k]["instruction_roofline_gcups"])
    out = {
        "metric": "all-pairs SW GCUPS", "value": gcups, "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        # the fp16 form holds the integer scores exactly (|H| <= 2048, rows of
        # at most 512 residues); longer rows run the packed int16 form
        "dtype": "f16 (exact integers) + int16" if t["kernel"] == 3 else "int16",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config],
                   "n_seqs": len(seqs), "unique_pairs": pairs, "all_pairs_cells": cells,
                   "computed_pairs": comp_pairs, "computed_cells": comp_cells,
                   "eps_max": args.eps_max, "eps_list_entries": nnz,
                   "kernel": {1: "register-profile int16", 2: "lds-profile int16",
                              3: "lds-profile exact fp16 (+int16 pass)"}.get(t["kernel"]),
                   "strip": t["strip"], "waves": t["waves"], "items": t["items"],
                   "autotune": tuning},
        "roofline": {"bound": "valu",
                     "kernel": ("sw_lds_mixed_kernel (one launch: exact fp16 form, int16 form for pairs "
                                "beyond it)" if mixed else
                                "sw_lds_f16_kernel (strips 64 and 32)" if "fp16" in forms else "sw_lds_kernel"),
                     "achieved": f16["gcups"], "peak": f16_peak,
                     "unit": "GCUPS (cells of the pairs the kernel computed / its time, HIP events on the "
                             "engine's stream; peak = the instruction roofline of its forms, each form's "
                             "cells at its own)",
                     "frac": f16["gcups"] / f16_peak,
                     "traffic": profiled_traffic(args.config, t["strip"]),
                     "step_traffic": profiled_traffic(args.config, t["strip"], "traffic_bytes_per_step"),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this workload, calibrated on a "
                                       "plain copy (profiles/); not measured in this run",
                     "step_kernel_ms": kernel_s * 1e3,
                     "step_kernel_gcups": comp_cells / kernel_s / 1e9,
                     # every pass of the step against its own form's roofline
                     "step_frac": at_roof_s / kernel_s,
                     "forms": forms,
                     # the fixed algorithmic count (6 int ops per cell) against the packed 16-bit
                     # issue peak of the chip: the fp16 form exceeds 1 here because it retires
                     # those 6 ops in 1.75 instructions per cell
                     "algorithmic_tops": OPS_PER_CELL * comp_cells / kernel_s / 1e12,
                     "algorithmic_peak_tops": PEAK_INT16_TOPS,
                     "algorithmic_ratio": OPS_PER_CELL * comp_cells / kernel_s / 1e12 / PEAK_INT16_TOPS,
                     "int16_only_gcups": comp_cells / int16_s / 1e9,
                     "int16_only_kernel_ms": int16_kernel_ms},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline: the reference's score() on a sample, one thread ...")
        out["cpu_baseline"] = cpu_baseline(seqs, args.cpu_sample)
        log(f"  {out['cpu_baseline']['value']:.3f} GCUPS; the clean-room port on the same sample ...")
        out["cpu_baseline_port"] = cpu_baseline(seqs, args.cpu_sample, port=True)
        out["cpu_baseline_port"]["port_over_reference"] = (out["cpu_baseline_port"]["value"]
                                                           / out["cpu_baseline"]["value"])
        # best effort on the host: every core this process may run on (at most 16)
        nt = max(1, min(16, len(os.sched_getaffinity(0))))
        log(f"  port {out['cpu_baseline_port']['value']:.3f} GCUPS; all {nt} cores ...")
        out["cpu_baseline_all_cores"] = cpu_baseline(seqs, 3 * args.cpu_sample, threads=nt)
        log("reference-faithful DBSCAN pass (families set sample) ...")
        out["cpu_baseline_reference_faithful"] = reference_faithful_dbscan(log)
    if not args.no_dbscan:
        log("adaptive DBSCAN CLI run ...")
        out["dbscan"] = run_dbscan(seqs, 100, 300, 5) if args.config == "tree20k" else \
            run_dbscan(seqs, 100, 100, 5)
        log(f"  {out['dbscan']['exact_s_per_run']:.2f} s")
    if not args.no_prefilter:
        log("learned prefilter vs exact DBSCAN (proposer100k workload) ...")
        pa = argparse.Namespace(**vars(args))
        pa.n, pa.eps = 0, 0
        res, workload, mfma, _ = proposer_eval(pa, pkg, log, "proposer100k")
        keys = ("eps", "minpts", "recall", "proposed_subset_of_exact", "precision", "candidates",
                "neighbour_pairs", "exact_neighbour_pairs", "exact_lists_s", "prefilter_lists_s",
                "exact_dbscan_s", "prefilter_dbscan_s", "speedup", "exact_all_pairs_estimate",
                "clusters_exact", "clusters_prefilter", "identical_clusters", "verify_gcups")
        block = {"workload": workload, **{k: res[k] for k in keys if k in res}}
        block["identical_partition"] = block.pop("identical_clusters")
        block["recall_bar"] = 0.999
        block["meets_recall_bar"] = bool(res["recall"] >= 0.999)
        block["pairwise_roofline"] = mfma
        out["dbscan_exact_vs_prefilter"] = block
    print(json.dumps(out), flush=True)


def reference_faithful_dbscan(log, n_sample=300, eps=400, minpts=5):
    """The reference's own cluster_dbscan_SW (CPU build, compiled from its
    sources into oracle/_ref) on the first n_sample sequences of the
    2,000-sequence 200-bp families set: one DBSCAN pass = n_sample region
    queries of n_sample score() calls each (dbscan.c:525-561), one thread,
    timed and extrapolated to N = 2,000 by N^2; the drop-in CLI is timed on
    the same sample (outputs compared) and on all 2,000 sequences."""
    ref = os.path.join(ROOT, "oracle", "_ref", "cluster_dbscan_SW")
    ours = os.path.join(PKG_DIR, "bin", "cluster_dbscan_SW")
    if not os.path.exists(ref):
        return None
    seqs = synth.families_200bp()
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for tag, data in (("sample", seqs[:n_sample]), ("full", seqs)):
            fa = os.path.join(tmp, tag + ".fasta")
            synth.write_fasta(fa, data)
            for who, exe in (("ref", ref), ("ours", ours)):
                if who == "ref" and tag == "full":
                    continue
                d = os.path.join(tmp, f"{tag}_{who}")
                os.makedirs(d)
                secs, so = _run_cli([exe, fa, str(eps), str(minpts), "o"], d)
                res[f"{tag}_{who}"] = (secs, so, _outputs(d, "o"))
    ref_s = res["sample_ref"][0]
    same = res["sample_ref"][1] == res["sample_ours"][1] and res["sample_ref"][2] == res["sample_ours"][2]
    est = ref_s * (len(seqs) / n_sample) ** 2
    log(f"  reference {n_sample} seqs: {ref_s:.1f} s (outputs identical to the drop-in: {same}); "
        f"extrapolated to {len(seqs)}: {est:.0f} s; drop-in on {len(seqs)}: {res['full_ours'][0]:.2f} s")
    return {"kind": "reference", "cores": 1, "unit": "s per DBSCAN pass",
            "value": est, "label": "extrapolated by N^2 from a timed sample",
            "sample": f"the reference's cluster_dbscan_SW (oracle/_ref, one thread) on the first {n_sample} "
                      f"of the 2,000 families_200bp sequences, eps {eps}, minpts {minpts}: {ref_s:.2f} s "
                      f"({n_sample}^2 score() calls); host: {cpu_model()}",
            "sample_s": ref_s, "sample_outputs_identical": bool(same),
            "drop_in_sample_s": res["sample_ours"][0], "drop_in_full_s": res["full_ours"][0],
            "speedup_full_vs_extrapolated": est / res["full_ours"][0]}


if __name__ == "__main__":
    main()
