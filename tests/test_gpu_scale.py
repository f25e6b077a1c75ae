"""GPU tests at the workloads' full sizes: the 105,000-sequence amplicon
lists (exact path), the learned proposer on 100,000 held-out sequences and
the 500,000-sequence heavy-tailed end-to-end run -- recall of the prefilter
lists against the exact lists, every prefilter entry an exact entry, and the
same DBSCAN partition (dbscan.c:950-1047 semantics) from both."""
import os
import sys

import numpy as np
import pytest

from conftest import PKG_DIR

pytestmark = pytest.mark.gpu

sys.path.insert(0, PKG_DIR)
import synth  # noqa: E402
from proposer_train import (lists_subset_and_recall, mix500k_workload,  # noqa: E402
                            proposer100k_workload, same_partition)


def _say(msg):
    print(msg, flush=True)


def _check_list_structure(rp, col, dist, n, eps):
    """Rows strictly ascending, no self entry, every d <= eps, and symmetric:
    (p, q, d) listed iff (q, p, d) is."""
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    assert np.all(col != rows)
    same_row = rows[1:] == rows[:-1]
    assert np.all(np.diff(col.astype(np.int64))[same_row] > 0)
    assert len(dist) == 0 or int(dist.max()) <= eps
    key = rows * n + col
    tkey = col.astype(np.int64) * n + rows
    order = np.argsort(tkey, kind="stable")
    assert np.array_equal(tkey[order], key)
    assert np.array_equal(dist[order], dist)


@pytest.mark.timeout(600)
def test_amplicon105k_eps_lists_at_full_size(pkg, oracle):
    """105,000 amplicon-like sequences (500 ancestors x 200 at 2-10% mutation
    plus 5% noise, lengths 150-300), exact lists at eps 300 over every unique
    pair: list structure on all rows; 400 random rows equal to one-vs-all
    launches (listed iff d <= eps, same d); 6 rows equal to the oracle."""
    seqs = synth.amplicon_families()
    n = len(seqs)
    eps = 300
    _say(f"amplicon105k: n={n}, eps {eps}")
    rng = np.random.default_rng(105)
    with pkg.Engine(seqs) as eng:
        rp, col, dist = eng.eps_csr(eps)
        t = eng.last_timing()
        _say(f"  lists: nnz {len(col)}, kernel {t['kernel_ms'] / 1e3:.2f} s, "
             f"{t['cells'] / t['kernel_ms'] / 1e6:.0f} GCUPS")
        assert len(col) > 10 * n
        _check_list_structure(rp, col, dist, n, eps)
        for p in rng.choice(n, 400, replace=False):
            d = eng.one_vs_all(int(p))
            want = np.flatnonzero(d <= eps)
            want = want[want != p]
            assert np.array_equal(col[rp[p]:rp[p + 1]], want), p
            assert np.array_equal(dist[rp[p]:rp[p + 1]], d[want].astype(np.uint16)), p
    allj = np.arange(n, dtype=np.int32)
    for p in rng.choice(n, 6, replace=False):
        others = allj[allj != p]
        ij = np.stack([np.full(n - 1, p, np.int32), others], 1)
        d = oracle.sw_pairs(seqs, ij, nthreads=16)
        keep = d <= eps
        assert np.array_equal(col[rp[p]:rp[p + 1]], others[keep]), p
        assert np.array_equal(dist[rp[p]:rp[p + 1]], d[keep].astype(np.uint16)), p


def _prefilter_vs_exact(pkg, enc, seqs, eps, minpts):
    with pkg.Engine(seqs) as eng:
        erp, ecol, edist = eng.eps_csr(eps)
        _say(f"  exact lists: nnz {len(ecol)}, kernel {eng.last_timing()['kernel_ms'] / 1e3:.2f} s")
        prp, pcol, pdist, st = enc.prefilter_csr(eng, eps)
        _say(f"  prefilter lists: nnz {len(pcol)}, candidates {st['candidates']}, "
             f"propose {st['propose_kernel_ms']:.1f} ms, verify {st['verify_kernel_ms']:.1f} ms")
    subset, recall = lists_subset_and_recall(prp, pcol, pdist, erp, ecol, edist)
    _say(f"  every prefilter entry exact: {subset}; recall {recall:.6f}")
    ecl = pkg.dbscan_csr(erp, ecol, edist, eps, eps, minpts)
    pcl = pkg.dbscan_csr(prp, pcol, pdist, eps, eps, minpts)
    _say(f"  DBSCAN: {len(ecl)} exact clusters, {len(pcl)} prefilter clusters")
    return subset, recall, same_partition(ecl, pcl), len(ecl)


@pytest.mark.timeout(900)
def test_proposer100k_recall_and_identical_partition(pkg):
    """The learned proposer trained on 2M exact-labelled pairs of the
    amplicon generator (20,000 ancestors x 20, seed 4) and evaluated on
    100,000 sequences of another draw at eps 300, minpts 5: prefilter lists
    are a subset of the exact lists with equal distances, recall >= 0.999,
    and DBSCAN over them gives the exact lists' partition."""
    enc, info, seqs, _ = proposer100k_workload(pkg, _say)
    try:
        subset, recall, same, k = _prefilter_vs_exact(pkg, enc, seqs, 300, 5)
    finally:
        enc.close()
    assert subset
    assert recall >= 0.999
    assert same and k > 100


@pytest.mark.timeout(900)
def test_mix500k_recall_and_identical_partition(pkg):
    """The end-to-end heavy-tailed workload: 500,000 sequences (lengths
    80-1500, Pareto 1.5; tree families, 171-bp tandem repeats, noise), encoder
    trained on four other draws, eps 200, minpts 5: prefilter lists are a
    subset of the exact lists with equal distances, recall >= 0.999, and
    DBSCAN over them gives the exact lists' partition."""
    enc, info, seqs, _ = mix500k_workload(pkg, _say)
    try:
        subset, recall, same, k = _prefilter_vs_exact(pkg, enc, seqs, 200, 5)
    finally:
        enc.close()
    assert subset
    assert recall >= 0.999
    assert same and k > 100
