"""CPU tests that pin the oracle (C restatement of smith-waterman.c and the
Python restatement of dbscan.c / cluster_io.c) to the compiled reference's
golden outputs, closed forms and brute force."""
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, cli_cases, read_case_files


# --------------------------------------------------------------------------
# Smith-Waterman distance
# --------------------------------------------------------------------------

def test_oracle_matches_reference_golden_pairs(oracle, golden_pairs):
    bad = [(a, b, d, oracle.sw_distance(a.encode(), b.encode()))
           for a, b, d in golden_pairs
           if oracle.sw_distance(a.encode(), b.encode()) != d]
    assert not bad, bad[:5]
    assert len(golden_pairs) > 1500


def test_oracle_batch_paths_match_single(oracle, golden_pairs):
    seqs = [p[0].encode() for p in golden_pairs[:200]] + [p[1].encode() for p in golden_pairs[:200]]
    ij = np.array([[k, k + 200] for k in range(200)], dtype=np.int32)
    got = oracle.sw_pairs(seqs, ij, nthreads=4)
    assert got.tolist() == [p[2] for p in golden_pairs[:200]]
    m = oracle.sw_matrix(seqs[:60], nthreads=4)
    for i in range(60):
        for j in range(60):
            assert m[i, j] == oracle.sw_distance(seqs[i], seqs[j])


def test_oracle_bruteforce_ac6(oracle, golden_bruteforce):
    strs = [s.encode() for s in golden_bruteforce["strings"]]
    want = np.array(golden_bruteforce["dist"], dtype=np.int64)
    got = oracle.sw_matrix(strs, nthreads=8)
    assert np.array_equal(got, want)


def test_oracle_invariants(oracle):
    rng = random.Random(5)
    for _ in range(400):
        alpha = rng.choice(["ACGT", "AC", "ACGTN"])
        a = "".join(rng.choice(alpha) for _ in range(rng.randint(0, 60))).encode()
        b = "".join(rng.choice(alpha) for _ in range(rng.randint(0, 60))).encode()
        d = oracle.sw_distance(a, b)
        assert d == oracle.sw_distance(b, a)                      # symmetry
        assert oracle.sw_distance(a, a) == 0                      # self distance
        assert 4 * abs(len(a) - len(b)) <= d <= 4 * max(len(a), len(b))
        if len(a) >= 1 and len(b) >= 1:
            # composition bound over x[1..] plus the terminator
            ca = np.bincount(np.frombuffer(a[1:], np.uint8), minlength=256)
            cb = np.bincount(np.frombuffer(b[1:], np.uint8), minlength=256)
            common = int(np.minimum(ca, cb).sum()) + 1
            assert d >= 4 * max(len(a), len(b)) - 4 * common
    # any two length-1 sequences: only the terminators are compared
    for x in b"ACGTN":
        for y in b"ACGTN":
            assert oracle.sw_distance(bytes([x]), bytes([y])) == 0


def _py_dp(a, b, skip0=True, terminator=True, gap=3, floor=True, opt=max):
    """Small pure-Python DP with switchable deviations from the reference."""
    A = a + b"\0" if terminator else a
    B = b + b"\0" if terminator else b
    i0 = 1 if skip0 else 0
    rows = range(i0, len(A))
    cols = range(i0, len(B))
    H = {}
    best = 0
    for i in rows:
        for j in cols:
            s = 4 if A[i] == B[j] else -4
            h = max(H.get((i - 1, j - 1), 0) + s, H.get((i - 1, j), 0) - gap,
                    H.get((i, j - 1), 0) - gap)
            if floor:
                h = max(h, 0)
            H[(i, j)] = h
            best = max(best, h)
    return 4 * opt(len(a), len(b)) - best


def test_golden_pairs_pin_each_quirk(golden_pairs):
    """The faithful DP reproduces the goldens and every plausible mistake
    (comparing index 0, dropping the terminator, a wrong gap, no floor, the
    wrong normaliser) fails at least one golden pair."""
    small = [(a.encode(), b.encode(), d) for a, b, d in golden_pairs
             if len(a) <= 24 and len(b) <= 24][:300]
    assert all(_py_dp(a, b) == d for a, b, d in small)
    variants = {
        "compare index 0": dict(skip0=False),
        "no terminator": dict(terminator=False),
        "gap 4": dict(gap=4),
        "no floor": dict(floor=False),
        "min-length normaliser": dict(opt=min),
    }
    for name, kw in variants.items():
        assert any(_py_dp(a, b, **kw) != d for a, b, d in small), name


# --------------------------------------------------------------------------
# FASTA reader / DBSCAN / adaptive / file formats against the reference CLI
# --------------------------------------------------------------------------

CASES = cli_cases()


@pytest.mark.parametrize("name,meta", CASES, ids=[c[0] for c in CASES])
def test_oracle_cli_matches_reference(oracle, name, meta):
    stdout, files = read_case_files(name, meta)
    seqs = oracle.read_fasta(os.path.join(GOLDEN, "inputs", meta["input"]))
    dist = oracle.sw_matrix(seqs, nthreads=8)
    if meta["program"] == "cluster_dbscan_SW":
        eps = float(np.float32(float(meta["args"][0])))
        minpts = int(meta["args"][1], 0)
        text, out, _ = oracle.cluster_dbscan_sw(seqs, eps, minpts, dist)
        got = {meta["prefix"] + k: v for k, v in out.items()}
    else:
        a = meta["args"]
        text, out, _ = oracle.adaptive_clustering(seqs, float(a[0]), float(a[1]),
                                                  int(a[2], 0), dist)
        got = {meta["prefix"] + k: v for k, v in out.items()}
    assert text.encode() == stdout
    assert sorted(got) == sorted(files)
    for k in files:
        assert got[k] == files[k], k


def _consumer_cases():
    d = os.path.join(GOLDEN, "consumers")
    for name in sorted(os.listdir(d)):
        meta = json.load(open(os.path.join(d, name, "cmd.json")))
        yield name, meta


@pytest.mark.parametrize("name,meta", [c for c in _consumer_cases() if "kmers10" not in c[0]],
                         ids=[c[0] for c in _consumer_cases() if "kmers10" not in c[0]])
def test_oracle_consumers_match_reference(oracle, name, meta):
    """The oracle's restatement of the other score() consumers reproduces the
    compiled reference's stdout and files (the 81M-pair wrap case runs on the
    GPU only)."""
    d = os.path.join(GOLDEN, "consumers", name)
    want = open(os.path.join(d, "stdout.txt"), "rb").read()
    argv = meta["argv"]
    inp = lambda k: oracle.read_fasta(os.path.join(GOLDEN, "inputs", argv[k][3:]))
    prog = meta["program"]
    if prog == "find_closest_sequence_SW":
        db, q = inp(0), inp(1)
        idx = int(argv[2])
        assert idx == 1 or len(q[idx - 1]) == len(q[0])
        text, f = oracle.find_closest(db, q[idx - 1])
        assert text.encode() == want
        assert f == open(os.path.join(d, argv[4]), "rb").read()
    elif prog == "find_satellite":
        got = oracle.find_satellite(inp(0), inp(1)[0], *[int(x) for x in argv[2:6]])
        assert got == want
    elif prog == "compareSW":
        assert oracle.compare_sw(inp(0), inp(1)).encode() == want
    elif prog == "compare_norms":
        proj = np.loadtxt(os.path.join(GOLDEN, "inputs", argv[1][3:]), ndmin=2)
        assert oracle.compare_norms(inp(0), proj[:, :int(argv[2])]).encode() == want
    else:
        assert oracle.silhouette(inp(0), inp(1)).encode() == want
