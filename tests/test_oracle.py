"""CPU tests that pin the oracle (C restatement of smith-waterman.c and the
Python restatement of dbscan.c / cluster_io.c) to the compiled reference's
golden outputs, closed forms and brute force."""
import json
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN, cli_cases, golden_arg, read_case_files, vector_cases


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


# --------------------------------------------------------------------------
# L1 / L2 vector tools (dbscan.c:563-780) against the compiled reference
# --------------------------------------------------------------------------

VEC_PROGRAMS = {"cluster_dbscan_pca": 2, "cluster_dbscan_kmerL1": 1, "cluster_dbscan_kmerL2": 2,
                "adaptive_clustering_PCA": 2, "adaptive_clustering_kmer_L1": 1,
                "adaptive_clustering_kmer_L2": 2}
VEC_CASES = vector_cases(VEC_PROGRAMS)


def vec_case_inputs(oracle, meta):
    """(sequences or None, values[dims][n], norm, eps, minpts) of a fixture,
    parsed as the reference's main() parses them."""
    prog, a = meta["program"], [golden_arg(x) for x in meta["argv"]]
    norm = VEC_PROGRAMS[prog]
    f32 = lambda t: float(oracle._strtof(t.encode()))
    if prog == "cluster_dbscan_pca":
        seqs = oracle.read_fasta(a[0])
        return seqs, oracle.read_projections(a[1], int(a[2]), len(seqs)), norm, f32(a[3]), int(a[4], 0)
    if prog.startswith("cluster_dbscan_kmer"):
        return oracle.read_fasta(a[0]), oracle.read_kmers(a[1]), norm, f32(a[2]), int(a[3], 0)
    if prog == "adaptive_clustering_PCA":
        seqs = oracle.read_fasta(a[0])
        return seqs, oracle.read_projections(a[7], int(a[6]), len(seqs)), norm, f32(a[1]), int(a[3], 0)
    return None, oracle.read_kmers(a[0]), norm, f32(a[1]), int(a[3], 0)


def vec_case_expected(name, meta):
    d = os.path.join(GOLDEN, "vectors", name)
    want = open(os.path.join(d, "stdout.txt"), "rb").read()
    files = {f: open(os.path.join(d, f), "rb").read() for f in meta["files"]}
    return want, files


@pytest.mark.parametrize("name,meta", VEC_CASES, ids=[c[0] for c in VEC_CASES])
def test_oracle_vector_tools_match_reference(oracle, name, meta):
    """The oracle's restatement of the L1 / L2 scans, DBSCAN and the tools'
    main() reproduces the compiled reference's stdout and every file."""
    want, files = vec_case_expected(name, meta)
    seqs, values, norm, eps, minpts = vec_case_inputs(oracle, meta)
    prog, argv = meta["program"], meta["argv"]
    if prog.startswith("cluster_dbscan"):
        with_values = prog == "cluster_dbscan_pca" and len(argv) == 7
        text, got, _ = oracle.cluster_dbscan_vec(seqs, values, norm, eps, minpts, with_values)
    else:
        a = [golden_arg(x) for x in argv]
        acc = oracle.vec_acc(values, norm)
        text, out, _ = oracle.adaptive_clustering(None, float(oracle._strtof(a[1].encode())),
                                                  float(oracle._strtof(a[2].encode())), minpts,
                                                  rq=oracle.region_query_from_acc(acc, norm),
                                                  n=values.shape[1])
        got = {"out" + k: v for k, v in out.items()}
    assert text.encode() == want
    assert sorted(got) == sorted(files)
    for k in files:
        assert got[k] == files[k], k


def _unfused_l2(values):
    """The L2 accumulator with the multiply rounded before the add (the
    reference's build without FMA), for the pinning test below."""
    v = values.astype(np.float32)
    n = v.shape[1]
    acc = np.zeros((n, n), np.float32)
    for k in range(v.shape[0]):
        d = (v[k][:, None] - v[k][None, :]).astype(np.float32)
        acc = (np.float32(1) * (d * d).astype(np.float32) + acc).astype(np.float32)
    return acc


@pytest.mark.parametrize("case,fused_in", [("pca_fma_in", True), ("pca_fma_out", False)])
def test_vector_fma_cases_pin_the_fused_step(oracle, case, fused_in):
    """The two-point fixtures separate the fused L2 step from the unfused
    one: the reference's cluster count agrees with the fused accumulator and
    not with the unfused one."""
    meta = dict(vector_cases())[case]
    seqs, values, norm, eps, minpts = vec_case_inputs(oracle, meta)
    thr = oracle.vec_threshold(2, eps)
    fused = oracle.vec_acc(values, 2)[0, 1]
    unfused = _unfused_l2(values)[0, 1]
    assert (fused <= thr) == fused_in and (unfused <= thr) != fused_in
    want, _ = vec_case_expected(case, meta)
    assert want == (b"1 clusters obtained \n" if fused_in else b"0 clusters obtained \n")


def test_vector_oracle_closed_forms(oracle):
    """Integer-valued vectors: every partial sum is an integer below 2^24, so
    the float32 accumulators equal the exact L1 / squared L2 distances."""
    rng = np.random.default_rng(3)
    for dims, n in ((1, 7), (3, 40), (64, 50), (300, 20)):
        x = rng.integers(-40, 40, size=(dims, n))
        d = x[:, :, None] - x[:, None, :]
        l1 = np.abs(d).sum(axis=0)
        l2 = (d * d).sum(axis=0)
        assert np.array_equal(oracle.vec_acc(x.astype(np.float32), 1), l1.astype(np.float32))
        assert np.array_equal(oracle.vec_acc(x.astype(np.float32), 2), l2.astype(np.float32))
        rows = rng.choice(n, size=min(n, 5), replace=False)
        assert np.array_equal(oracle.vec_acc(x.astype(np.float32), 2, rows), l2[rows].astype(np.float32))
    # non-integer data: within float32 rounding of the float64 distances, with
    # the difference taken as query minus candidate
    x = rng.normal(size=(5, 30)).astype(np.float32)
    d = x[:, :, None].astype(np.float64) - x[:, None, :].astype(np.float64)
    assert np.allclose(oracle.vec_acc(x, 2), (d * d).sum(axis=0), rtol=1e-5, atol=1e-6)
    assert np.allclose(oracle.vec_acc(x, 1), np.abs(d).sum(axis=0), rtol=1e-5, atol=1e-6)
    # NaN / inf inputs: the comparisons fail, the point is nobody's neighbour
    y = x.copy()
    y[2, 4] = np.nan
    y[0, 7] = np.inf
    acc = oracle.vec_acc(y, 2)
    assert np.isnan(acc[4]).all() and np.isnan(acc[:, 4]).all() and np.isnan(acc[7, 7])


def test_vector_oracle_parsers(oracle):
    """k-mer shape / values and projection parsing against the fasta2kmer
    format (kmers.c:378-389) and numpy."""
    path = os.path.join(GOLDEN, "inputs", "c2_small_k4.kmer")
    data = open(path, "rb").read()
    assert oracle.kmer_shape(data) == (256, 256)
    rows = [ln.split(b"\t") for ln in data.decode().encode().split(b"\n") if ln]
    v = oracle.read_kmers(path)
    assert np.array_equal(v, np.array([[float(t) for t in r[1:]] for r in rows], np.float32).T)
    p = oracle.read_projections(os.path.join(GOLDEN, "inputs", "c1_small_proj5.txt"), 5, 320)
    ref = np.loadtxt(os.path.join(GOLDEN, "inputs", "c1_small_proj5.txt"), dtype=np.float64)
    assert np.array_equal(p, ref.astype(np.float32).T)


KMER_CASES = vector_cases({"fasta2kmer"})


@pytest.mark.parametrize("name,meta", KMER_CASES, ids=[c[0] for c in KMER_CASES])
def test_oracle_fasta2kmer_matches_reference(oracle, name, meta):
    """The oracle's k-mer counting (kmers.c:159-243 with its quirks: the last
    window not counted, wrapping short counts, column orders) reproduces the
    compiled reference's fasta2kmer output."""
    a = meta["argv"]
    seqs = oracle.read_fasta(golden_arg(a[0]))
    want, _ = vec_case_expected(name, meta)
    assert oracle.kmer_text(oracle.kmer_frequencies(seqs, int(a[1]), bool(int(a[3])))) == want


def test_oracle_kmer_inputs_are_reference_outputs(oracle):
    """The k-mer inputs of the vector fixtures were written by the reference's
    fasta2kmer; the oracle reproduces them too."""
    for fasta, k, kmer in (("c1_small.fasta", 3, "c1_small_k3.kmer"), ("c2_small.fasta", 4, "c2_small_k4.kmer")):
        seqs = oracle.read_fasta(os.path.join(GOLDEN, "inputs", fasta))
        want = open(os.path.join(GOLDEN, "inputs", kmer), "rb").read()
        assert oracle.kmer_text(oracle.kmer_frequencies(seqs, k, False)) == want


def test_kmer_oracle_closed_forms(oracle):
    """Brute force on tiny inputs: row sums are the number of valid windows
    below len - k, DNA column names decode little-endian, protein names are
    lexicographic in the reference's alphabet."""
    names = oracle.kmer_names(2, False)
    assert names[:5] == [b"AA", b"CA", b"GA", b"TA", b"AC"]
    assert oracle.kmer_names(2, True)[:3] == [b"AA", b"AC", b"AD"] and oracle.kmer_names(1, True)[-1] == b"Y"
    rng = np.random.default_rng(9)
    for _ in range(30):
        s = bytes(rng.choice(list(b"ACGTN"), size=int(rng.integers(0, 40))).astype(np.uint8))
        for k in (1, 2, 3):
            f = oracle.kmer_frequencies([s], k, False)[0]
            valid = sum(1 for i in range(max(0, len(s) - k)) if b"N" not in s[i:i + k])
            assert int(f.astype(np.int64).sum()) == valid
