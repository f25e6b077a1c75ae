#!/usr/bin/env python3
"""Benchmark: all-pairs Smith-Waterman GCUPS on one MI355X.

Workload (one step = one pass of the hot path over the whole batch):
  uneven-length set of 20,000 sequences from a 3-level evolution tree
  (100 roots with lognormal lengths, median 300, clipped to [100, 600];
  10 children at 8% mutation; 20 leaves per child at 3%; sub:indel 4:1),
  all 2.0e8 unique pairs through the length-bucketed persistent gfx950
  kernel, with the epsilon-neighbour lists (d <= eps_max) emitted on the
  device.  Sequences are uploaded once before timing (resident in HBM).

GCUPS = sum over unique pairs of la*lb / step seconds / 1e9.

Also reported:
  roofline     -- the SW kernel's integer-op rate (6 ops per cell, measured
                  with HIP events on the engine's stream) against the packed
                  int16 VALU peak of the chip;
  cpu_baseline -- the reference's own score() (compiled from its source into
                  oracle/_ref/) on a bounded random sample of the same
                  sequences, single thread, as the reference's CPU path runs;
  dbscan       -- wall seconds of the drop-in adaptive_clustering_SW CLI on the
                  same 20,000 sequences (FASTA in, split-set files out).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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

WORKLOADS = {
    "c2": "uneven-length all-pairs SW: 20,000 seqs, 3-level tree, lognormal lengths "
          "median 300 in [100,600]",
    "c3": "amplicon-scale all-pairs SW: 105,000 seqs (500 ancestors x 200 at 2-10% "
          "mutation + 5% noise), lengths uniform in [150,300]",
}

# MI355X: 256 CUs x 64 lanes x 2 (packed int16) x 2.4 GHz
PEAK_INT16_TOPS = 256 * 64 * 2 * 2.4e9 / 1e12
OPS_PER_CELL = 6


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


def cpu_baseline(seqs, sample_n, seed=7):
    """Reference score() (oracle/_ref/libref_sw.so) over all unique pairs of
    a random sample, one thread; the oracle's C restatement if the reference
    build is absent."""
    rng = random.Random(seed)
    sample = rng.sample(range(len(seqs)), sample_n)
    sub = [seqs[i] for i in sample]
    bufs = [ctypes.create_string_buffer(s, len(s) + 1) for s in sub]
    ptrs = (ctypes.c_char_p * len(sub))(*[ctypes.cast(b, ctypes.c_char_p) for b in bufs])
    lens = (ctypes.c_size_t * len(sub))(*[len(s) for s in sub])
    ref = os.path.join(ROOT, "oracle", "_ref", "libref_sw.so")
    if os.path.exists(ref):
        lib, fn, kind = ctypes.CDLL(ref), "ref_cpu_allpairs", "reference"
    else:
        lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle_sw.so"))
        fn, kind = "oracle_cpu_allpairs", "port"
    f = getattr(lib, fn)
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
    cs = ctypes.c_uint64(0)
    secs = f(ptrs, lens, len(sub), 1, 0, None, ctypes.byref(cs))
    L = np.array([len(s) for s in sub], dtype=np.float64)
    cells = 0.5 * (L.sum() ** 2 - (L * L).sum())
    return {"value": cells / secs / 1e9, "unit": "GCUPS", "cores": 1, "kind": kind,
            "sample": f"all {sample_n * (sample_n - 1) // 2} unique pairs of {sample_n} "
                      f"random sequences of the workload ({cells:.3g} cells, {secs:.1f} s); "
                      f"host: {cpu_model()}, nproc {os.cpu_count()}"}


def profiled_traffic(config):
    """HBM bytes per step of the SW kernels, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    plain copy kernel); None if there is no summary for it."""
    path = os.path.join(ROOT, "profiles", {"c2": "sw_lds_f16_kernel_config2_20k.md"}.get(config, "-"))
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_step:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps-max", type=int, default=2500)
    ap.add_argument("--n", type=int, default=0, help="use only the first n sequences")
    ap.add_argument("--cpu-sample", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-dbscan", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--config", default="c2", choices=["c2", "c3"],
                    help="c2: 20,000 uneven-length (headline); c3: 105,000 amplicon-scale")
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

    pkg = load_pkg()
    print("generating workload ...", file=sys.stderr, flush=True)
    seqs = synth.config2() if args.config == "c2" else synth.config3()
    if args.n:
        seqs = seqs[:args.n]
    L = np.array([len(s) for s in seqs], dtype=np.float64)
    cells = 0.5 * (L.sum() ** 2 - (L * L).sum())
    pairs = len(seqs) * (len(seqs) - 1) // 2

    eng = pkg.Engine(seqs)
    tuning = None
    if not args.no_autotune:
        tuning = eng.autotune(use_cache=True)
        print(f"autotune: {tuning}", file=sys.stderr, flush=True)
    print(f"n={len(seqs)} pairs={pairs} cells={cells:.3e}; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        eng.eps_count(args.eps_max)
    times, kms = [], []
    nnz = 0
    for k in range(args.steps):
        t0 = time.perf_counter()
        nnz = eng.eps_count(args.eps_max)
        times.append(time.perf_counter() - t0)
        t = eng.last_timing()
        kms.append(t["kernel_ms"])
        print(f"step {k}: {times[-1] * 1e3:.1f} ms (kernel {kms[-1]:.1f} ms), "
              f"{cells / times[-1] / 1e9:.1f} GCUPS, nnz {nnz}", file=sys.stderr, flush=True)
    t = eng.last_timing()
    # the same step with the int16 form of the LDS kernel only (integer VALU path)
    eng.set_tuning(2, 32, 16)
    t0 = time.perf_counter()
    eng.eps_count(args.eps_max)
    int16_s = time.perf_counter() - t0
    int16_kernel_ms = eng.last_timing()["kernel_ms"]
    print(f"int16-only step: {int16_s * 1e3:.1f} ms, {cells / int16_s / 1e9:.1f} GCUPS",
          file=sys.stderr, flush=True)
    eng.close()
    step_s = float(np.mean(times))
    kernel_s = float(np.mean(kms)) / 1e3
    gcups = cells / step_s / 1e9
    achieved = OPS_PER_CELL * cells / kernel_s / 1e12
    out = {
        "metric": "all-pairs SW GCUPS", "value": gcups, "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int16",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config],
                   "n_seqs": len(seqs), "unique_pairs": pairs, "cells": cells,
                   "eps_max": args.eps_max, "eps_list_entries": nnz,
                   "kernel": {1: "register-profile int16", 2: "lds-profile int16",
                              3: "lds-profile exact fp16 (+int16 pass)"}.get(t["kernel"]),
                   "strip": t["strip"], "waves": t["waves"], "items": t["items"],
                   "autotune": tuning},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT16_TOPS,
                     "unit": "TOP/s (packed 16-bit VALU; 6 algorithmic ops per cell)",
                     "frac": achieved / PEAK_INT16_TOPS,
                     "traffic": profiled_traffic(args.config),
                     "kernel": "sw_lds_f16_kernel + sw_lds_kernel" if t["kernel"] == 3
                               else {1: "sw_tiles_kernel", 2: "sw_lds_kernel"}.get(t["kernel"]),
                     "kernel_ms": kernel_s * 1e3,
                     "kernel_gcups": cells / kernel_s / 1e9,
                     "int16_only_gcups": cells / int16_s / 1e9,
                     "int16_only_kernel_ms": int16_kernel_ms},
    }
    if not args.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(seqs, args.cpu_sample)
    if not args.no_dbscan and args.config == "c2":
        print("adaptive DBSCAN run ...", file=sys.stderr, flush=True)
        out["dbscan"] = run_dbscan(seqs, 100, 300, 5)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
