"""GPU parity tests: the gfx950 engine (through its C ABI) against the
reference's golden vectors and the oracle, bit-exact."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(GOLDEN), "..", "mnhn-tree-tools_amd"))
import synth  # noqa: E402


def _say(msg):
    print(msg, flush=True)


def _rand_set(rng, n, lo, hi, alphabet=b"ACGT"):
    alpha = np.frombuffer(alphabet, np.uint8)
    return [alpha[rng.integers(0, len(alpha), rng.integers(lo, hi + 1))].tobytes()
            for _ in range(n)]


@pytest.mark.parametrize("strip", [16, 32])
def test_golden_pairs_acgt(pkg, golden_pairs, strip):
    pairs = [(a.encode(), b.encode(), d) for a, b, d in golden_pairs
             if set(a + b) <= set("ACGT")]
    seqs = [p[0] for p in pairs] + [p[1] for p in pairs]
    _say(f"golden ACGT pairs: {len(pairs)} strip {strip}")
    with pkg.Engine(seqs) as eng:
        assert eng.alphabet == 4
        eng.set_tuning(strip=strip)
        ij = np.array([[k, k + len(pairs)] for k in range(len(pairs))], np.int32)
        got = eng.pairs(ij)
        got_rev = eng.pairs(ij[:, ::-1])
    want = np.array([p[2] for p in pairs])
    assert np.array_equal(got, want)
    assert np.array_equal(got_rev, want)


@pytest.mark.parametrize("strip", [16, 32])
def test_golden_pairs_all_alphabets(pkg, golden_pairs, strip):
    pairs = [(a.encode(), b.encode(), d) for a, b, d in golden_pairs]
    seqs = [p[0] for p in pairs] + [p[1] for p in pairs]
    _say(f"golden pairs (generic alphabet): {len(pairs)} strip {strip}")
    with pkg.Engine(seqs) as eng:
        assert eng.alphabet == 0
        eng.set_tuning(strip=strip)
        ij = np.array([[k, k + len(pairs)] for k in range(len(pairs))], np.int32)
        got = eng.pairs(ij)
    assert np.array_equal(got, np.array([p[2] for p in pairs]))


def test_bruteforce_ac6_matrix(pkg, golden_bruteforce):
    strs = [s.encode() for s in golden_bruteforce["strings"]]
    want = np.array(golden_bruteforce["dist"], dtype=np.int32)
    _say(f"bruteforce matrix n={len(strs)}")
    with pkg.Engine(strs) as eng:
        got = eng.matrix()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("alphabet", [b"ACGT", b"ACGTN"])
@pytest.mark.parametrize("strip", [16, 32])
def test_matrix_vs_oracle_mixed_lengths(pkg, oracle, alphabet, strip):
    rng = np.random.default_rng(7 + strip)
    seqs = _rand_set(rng, 260, 0, 700, alphabet)
    # related sequences so that long local alignments occur
    base = _rand_set(rng, 4, 300, 700, b"ACGT")
    for b in base:
        for _ in range(10):
            seqs.append(synth.to_bytes(synth.mutate(
                rng, np.frombuffer(b, np.uint8).copy() % 4, 0.04, 0.01)))
    _say(f"matrix vs oracle n={len(seqs)} alphabet={alphabet} strip={strip}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    with pkg.Engine(seqs) as eng:
        eng.set_tuning(strip=strip)
        got = eng.matrix()
        row = eng.one_vs_all(17)
        row2 = eng.one_vs_all(len(seqs) - 1)
    assert np.array_equal(got, want)
    assert np.array_equal(row, want[17])
    assert np.array_equal(row2, want[-1])


def test_pairs_random_order_and_duplicates(pkg, oracle):
    rng = np.random.default_rng(11)
    seqs = _rand_set(rng, 500, 1, 400)
    ij = rng.integers(0, len(seqs), size=(3000, 2)).astype(np.int32)
    ij[:50, 1] = ij[:50, 0]  # self pairs
    _say("pairs: 3000 random pairs")
    with pkg.Engine(seqs) as eng:
        got = eng.pairs(ij)
    want = oracle.sw_pairs(seqs, ij, nthreads=16)
    assert np.array_equal(got, want)
    assert np.all(got[:50] == 0)


def _csr_to_dense(rp, col, dist, n):
    m = np.full((n, n), -1, dtype=np.int32)
    for i in range(n):
        m[i, col[rp[i]:rp[i + 1]]] = dist[rp[i]:rp[i + 1]]
    return m


@pytest.mark.parametrize("eps", [0, 150, 400, 100000])
def test_eps_csr_vs_oracle(pkg, oracle, eps):
    seqs = synth.families_200bp(seed=21, n_anc=6, per_anc=50)
    _say(f"eps csr n={len(seqs)} eps={eps}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    with pkg.Engine(seqs) as eng:
        rp, col, dist = eng.eps_csr(eps)
    n = len(seqs)
    assert rp[0] == 0 and rp[-1] == len(col)
    for i in range(n):
        c = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(c) > 0)          # ascending, no duplicates
        assert not np.any(c == i)              # self not stored
    dense = _csr_to_dense(rp, col, dist, n)
    expect = np.where(want <= eps, want, -1)
    np.fill_diagonal(expect, -1)
    assert np.array_equal(dense, expect)


@pytest.mark.parametrize("eps", [0, 40, 250, 900])
def test_eps_csr_length_pruning_is_exact(pkg, oracle, monkeypatch, eps):
    """The epsilon pass leaves out pairs with match * |la - lb| > eps; over
    widely spread lengths the lists must equal the oracle's with and without
    that pruning, and the pruned pass must compute fewer pairs."""
    rng = np.random.default_rng(100 + eps)
    seqs = _rand_set(rng, 300, 0, 700)
    base = _rand_set(rng, 6, 50, 650, b"ACGT")
    for b in base:
        for _ in range(12):
            seqs.append(synth.to_bytes(synth.mutate(
                rng, np.frombuffer(b, np.uint8).copy() % 4, 0.03, 0.01)))
    n = len(seqs)
    _say(f"eps csr length pruning n={n} eps={eps}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    expect = np.where(want <= eps, want, -1)
    np.fill_diagonal(expect, -1)
    got = {}
    with pkg.Engine(seqs) as eng:
        for prune in ("0", "1"):
            monkeypatch.setenv("MNHN_SW_LENGTH_PRUNE", prune)
            rp, col, dist = eng.eps_csr(eps)
            got[prune] = (eng.last_timing()["pairs"], _csr_to_dense(rp, col, dist, n))
    assert got["0"][0] == n * (n - 1) // 2
    assert got["1"][0] < got["0"][0]
    assert np.array_equal(got["0"][1], expect)
    assert np.array_equal(got["1"][1], expect)


@pytest.mark.parametrize("strip", [None, 32])
def test_eps_csr_one_launch_for_both_forms(pkg, oracle, monkeypatch, strip):
    """Lengths on both sides of 512 residues put items of the exact fp16 form
    and of the int16 form in one epsilon pass: by default they run as one
    launch (the form per item), with MNHN_SW_MIXED=0 as a launch per form;
    both give the oracle's lists, and the timing reports the one launch with
    the cells of each form."""
    rng = np.random.default_rng(77)
    seqs = _rand_set(rng, 120, 300, 800)
    base = _rand_set(rng, 4, 400, 750, b"ACGT")
    for b in base:
        for _ in range(15):
            seqs.append(synth.to_bytes(synth.mutate(
                rng, np.frombuffer(b, np.uint8).copy() % 4, 0.03, 0.01)))
    n, eps = len(seqs), 1200
    _say(f"eps csr one launch n={n} eps={eps} strip={strip}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    expect = np.where(want <= eps, want, -1)
    np.fill_diagonal(expect, -1)
    with pkg.Engine(seqs) as eng:
        if strip:
            eng.set_tuning(3, strip, 0)
        for mixed in ("1", "0"):
            monkeypatch.setenv("MNHN_SW_MIXED", mixed)
            rp, col, dist = eng.eps_csr(eps)
            t = eng.last_timing()
            assert np.array_equal(_csr_to_dense(rp, col, dist, n), expect)
            assert t["mixed_launch"] == (1 if mixed == "1" else 0)
            assert t["form_cells"][0] > 0 and t["form_cells"][1] > 0
            if mixed == "1":
                assert t["form_ms"][1] == 0 and t["form_ms"][0] > 0


def test_families_200bp_full_matrix_and_eps(pkg, oracle):
    seqs = synth.families_200bp(seed=1)
    n = len(seqs)
    _say(f"families_200bp full n={n}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    with pkg.Engine(seqs) as eng:
        got = eng.matrix()
        rp, col, dist = eng.eps_csr(400)
        t = eng.last_timing()
    assert np.array_equal(got, want)
    dense = _csr_to_dense(rp, col, dist, n)
    expect = np.where(want <= 400, want, -1)
    np.fill_diagonal(expect, -1)
    assert np.array_equal(dense, expect)
    _say(f"families_200bp eps kernel {t['kernel_ms']:.2f} ms, {t['cells'] / t['kernel_ms'] / 1e6:.1f} GCUPS")


def test_evolution_tree_full_size_properties(pkg, oracle):
    """Full 20,000-sequence uneven-length set: size-independent properties
    plus oracle checks on sampled pairs (neighbours and non-neighbours)."""
    seqs = synth.evolution_tree()
    n = len(seqs)
    eps = 600
    _say(f"evolution_tree full n={n}")
    with pkg.Engine(seqs) as eng:
        rp, col, dist = eng.eps_csr(eps)
        t = eng.last_timing()
        _say(f"evolution_tree eps kernel {t['kernel_ms']:.1f} ms, {t['cells'] / t['kernel_ms'] / 1e6:.1f} GCUPS, nnz={len(col)}")
        rng = np.random.default_rng(3)
        ij = rng.integers(0, n, size=(4000, 2)).astype(np.int32)
        d_pairs = eng.pairs(ij)
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(dist <= eps) and np.all(dist >= 0)
    assert not np.any(rows == col)
    # symmetric: (i, j, d) present iff (j, i, d) present
    key = rows.astype(np.int64) * n + col
    key_t = col.astype(np.int64) * n + rows
    o1, o2 = np.argsort(key), np.argsort(key_t)
    assert np.array_equal(key[o1], key_t[o2])
    assert np.array_equal(dist[o1], dist[o2])
    # sampled CSR entries are exact
    pick = rng.choice(len(col), size=min(3000, len(col)), replace=False)
    ij2 = np.stack([rows[pick], col[pick]], 1).astype(np.int32)
    assert np.array_equal(dist[pick], oracle.sw_pairs(seqs, ij2, nthreads=16))
    # random pairs: exact, and present in the CSR iff d <= eps
    want = oracle.sw_pairs(seqs, ij, nthreads=16)
    assert np.array_equal(d_pairs, want)
    keyset = set(key.tolist())
    for (i, j), d in zip(ij.tolist(), want.tolist()):
        if i != j:
            assert ((i * n + j) in keyset) == (d <= eps)


def test_autotune_choice_is_exact_and_cached(pkg, oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("MNHN_SW_TUNE_CACHE", str(tmp_path / "tune.jsonl"))
    seqs = synth.families_200bp(seed=31, n_anc=8, per_anc=60)
    _say(f"autotune n={len(seqs)}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    with pkg.Engine(seqs) as eng:
        t1 = eng.autotune(use_cache=True)
        assert t1["from_cache"] == 0 and t1["kernel"] in (1, 2, 3) and t1["gcups"] > 0
        got = eng.matrix()
        assert eng.last_timing()["strip"] == t1["strip"]
    assert np.array_equal(got, want)
    with pkg.Engine(seqs) as eng:
        t2 = eng.autotune(use_cache=True)
        assert t2["from_cache"] == 1
        assert (t2["kernel"], t2["strip"], t2["waves_per_cu"]) == (t1["kernel"], t1["strip"], t1["waves_per_cu"])
        assert np.array_equal(eng.matrix(), want)


@pytest.mark.parametrize("kernel,strip,wpc", [(1, 16, 0), (1, 32, 4), (2, 16, 24), (2, 32, 12), (2, 64, 8), (3, 16, 16), (3, 32, 12), (3, 64, 8)])
def test_every_tuning_is_exact(pkg, oracle, kernel, strip, wpc):
    rng = np.random.default_rng(kernel * 100 + strip)
    seqs = _rand_set(rng, 200, 0, 520)
    _say(f"tuning kernel={kernel} strip={strip} wpc={wpc}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    with pkg.Engine(seqs) as eng:
        eng.set_tuning(kernel, strip, wpc)
        got = eng.matrix()
        row = eng.one_vs_all(5)
        t = eng.last_timing()
    assert (t["kernel"], t["strip"]) == (kernel, strip)
    assert np.array_equal(got, want)
    assert np.array_equal(row, want[5])


@pytest.mark.parametrize("tuning", [None, (3, 64, 8), (2, 16, 16)])
def test_block_vs_oracle(pkg, oracle, tuning):
    """Rectangular blocks with repeated and overlapping rows/columns, ragged
    lengths (incl. empty sequences), and a column count that is no multiple
    of 128."""
    rng = np.random.default_rng(61)
    seqs = _rand_set(rng, 240, 0, 600)
    base = _rand_set(rng, 3, 80, 500, b"ACGT")
    for b in base:
        for _ in range(12):
            seqs.append(synth.to_bytes(synth.mutate(
                rng, np.frombuffer(b, np.uint8).copy() % 4, 0.04, 0.01)))
    n = len(seqs)
    want = oracle.sw_matrix(seqs, nthreads=16)
    rows = np.concatenate([rng.choice(n, 90), [0, 0, n - 1]]).astype(np.int32)
    cols = np.concatenate([rng.choice(n, 200), rows[:7]]).astype(np.int32)
    _say(f"block {len(rows)} x {len(cols)} tuning={tuning}")
    with pkg.Engine(seqs) as eng:
        if tuning:
            eng.set_tuning(*tuning)
        got = eng.block(rows, cols)
        one = eng.block(rows[:1], np.arange(n, dtype=np.int32))
        empty = eng.block(rows[:0], cols)
    assert np.array_equal(got, want[np.ix_(rows, cols)])
    assert np.array_equal(one[0], want[rows[0]])
    assert empty.shape == (0, len(cols))


@pytest.mark.parametrize("eps", [60, 400])
def test_eps_csr_radix_form_equals_rank_form(pkg, oracle, monkeypatch, eps):
    """The radix-sort key -> CSR form (used when n is too large for the
    row-ranking form's LDS bitmap; forced here with MNHN_CSR_SORT=radix)
    gives the same lists, and both equal the oracle's."""
    rng = np.random.default_rng(7 + eps)
    seqs = _rand_set(rng, 200, 0, 400)
    for b in _rand_set(rng, 8, 60, 300):
        for _ in range(25):
            seqs.append(synth.to_bytes(synth.mutate(rng, np.frombuffer(b, np.uint8).copy() % 4, 0.04, 0.01)))
    n = len(seqs)
    _say(f"eps csr radix form n={n} eps={eps}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    expect = np.where(want <= eps, want, -1)
    np.fill_diagonal(expect, -1)
    with pkg.Engine(seqs) as eng:
        monkeypatch.delenv("MNHN_CSR_SORT", raising=False)
        a = eng.eps_csr(eps)
        monkeypatch.setenv("MNHN_CSR_SORT", "radix")
        b = eng.eps_csr(eps)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(_csr_to_dense(*b, n), expect)


def test_eps_csr_beyond_row_bitmap_limit(pkg, oracle):
    """700,000 short sequences (more than the 651,000 columns the row-ranking
    form can hold in LDS): the radix form is taken; rows are sorted and 40
    sampled rows equal the oracle's."""
    rng = np.random.default_rng(70)
    n_anc, per = 20000, 35
    letters = np.frombuffer(b"ACGT", np.uint8)
    seqs = []
    for _ in range(n_anc):
        a = rng.integers(0, 4, int(rng.integers(10, 21)))
        copies = np.tile(a, (per, 1))
        hit = rng.random(copies.shape) < 0.05
        copies[hit] = rng.integers(0, 4, int(hit.sum()))
        seqs.extend(letters[row].tobytes() for row in copies)
    n = len(seqs)
    eps = 12
    _say(f"eps csr beyond bitmap limit n={n} eps={eps}")
    with pkg.Engine(seqs) as eng:
        rp, col, dist = eng.eps_csr(eps)
    assert len(rp) == n + 1 and rp[-1] == len(col)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    key = rows * n + col
    assert np.all(np.diff(key) > 0)
    assert np.all(dist <= eps)
    sample = rng.choice(n, 40, replace=False)
    all_j = np.arange(n, dtype=np.int32)
    ij = np.concatenate([np.stack([np.full(n - 1, p, np.int32), all_j[all_j != p]], 1) for p in sample])
    d = oracle.sw_pairs(seqs, ij, nthreads=16).reshape(len(sample), n - 1)
    for k, p in enumerate(sample):
        others = all_j[all_j != p]
        want = others[d[k] <= eps]
        assert np.array_equal(col[rp[p]:rp[p + 1]], want), p
        assert np.array_equal(dist[rp[p]:rp[p + 1]], d[k][d[k] <= eps].astype(np.uint16)), p


def test_eps_key_buffer_sized_from_sample(pkg, monkeypatch):
    """A first epsilon pass whose keys may exceed the starting buffer sizes it
    from a sample of the work (here the starting buffer holds 1,250 keys):
    the lists equal those of a context that starts with room for every
    pair, and a later, larger pass grows the buffer as needed."""
    rng = np.random.default_rng(91)
    seqs = _rand_set(rng, 3000, 60, 140)
    with pkg.Engine(seqs) as eng:
        want = eng.eps_csr(500)
        want_wide = eng.eps_csr(700)
    monkeypatch.setenv("MNHN_SW_KEY_BUDGET_GB", "0.00001")
    with pkg.Engine(seqs) as eng:
        got = eng.eps_csr(500)
        got_wide = eng.eps_csr(700)
    _say(f"sampled key buffer: nnz {len(got[1])} and {len(got_wide[1])}")
    assert len(want[1]) > 100000
    for g, w in ((got, want), (got_wide, want_wide)):
        assert all(np.array_equal(a, b) for a, b in zip(g, w))


def _long_set(rng, alphabet):
    """Sequences of 9,000-17,200 residues: mutated copies of one ancestor
    (related pairs score far above 2^16), an unrelated one and a short one."""
    alpha = np.frombuffer(alphabet, np.uint8)
    anc = alpha[rng.integers(0, len(alpha), 17200)]
    seqs = []
    for L, p in ((17200, 0.02), (16700, 0.05), (12000, 0.04), (9000, 0.02)):
        s = anc[:L].copy()
        flip = rng.random(L) < p
        s[flip] = alpha[rng.integers(0, len(alpha), int(flip.sum()))]
        seqs.append(s.tobytes())
    seqs.append(alpha[rng.integers(0, len(alpha), 16500)].tobytes())
    seqs.append(anc[5000:5300].tobytes())
    return seqs


@pytest.mark.parametrize("alphabet", [b"ACGT", b"ACGTN"])
def test_long_sequences_run_the_32bit_path(pkg, oracle, alphabet):
    """Pairs of sequences longer than 16,383 residues (match * min(la, lb) >
    65535, beyond the 16-bit lanes) run the 32-bit kernel: the matrix, pairs,
    a one-vs-all row and the epsilon lists up to 65535 equal the oracle's
    (distances up to 68,800), epsilon lists beyond 65535 are refused with
    MNHN_ERR_RANGE, and DBSCAN passes over the lists equal the oracle's."""
    rng = np.random.default_rng(77 + len(alphabet))
    seqs = _long_set(rng, alphabet)
    n = len(seqs)
    _say(f"long sequences ({alphabet.decode()}): lengths {[len(s) for s in seqs]}")
    want = oracle.sw_matrix(seqs, nthreads=16)
    assert want.max() > 65535 and (want[want > 0] < 40000).any()
    with pkg.Engine(seqs) as eng:
        assert eng.alphabet == (4 if alphabet == b"ACGT" else 0)
        got = eng.matrix()
        assert eng.last_timing()["lane_bits"] == 32
        assert np.array_equal(got, want)
        rp, col, dist = eng.eps_csr(65535)
        for p in range(n):
            others = np.array([j for j in range(n) if j != p])
            keep = want[p, others] <= 65535
            assert np.array_equal(col[rp[p]:rp[p + 1]], others[keep]), p
            assert np.array_equal(dist[rp[p]:rp[p + 1]], want[p, others][keep].astype(np.uint16)), p
        for e in (30000.5, 40000, 65535):
            want_cl = oracle.dbscan(n, e, 1, oracle.region_query_from_matrix(want))
            assert pkg.dbscan_csr(rp, col, dist, 65535, e, 1) == [c.members for c in want_cl.clusters], e
        with pytest.raises(pkg.MnhnError):
            eng.eps_csr(70000)
        if alphabet == b"ACGT":
            ij = np.array([[0, 1], [4, 1], [5, 0], [2, 3], [1, 1]], np.int32)
            assert np.array_equal(eng.pairs(ij), want[ij[:, 0], ij[:, 1]])
            assert np.array_equal(eng.one_vs_all(1), want[1])
