"""CPU tests of the proposer's threshold calibration (proposer_train.py): the
length buckets and the per-bucket thresholds that keep the calibration pairs."""
import numpy as np

import proposer_train as pt


def _pairs(rng):
    """Synthetic calibration pairs: many short ones, a few long ones, with
    cosines and radius excesses that tighten with length."""
    lmax = np.concatenate([rng.integers(90, 150, 50_000), rng.integers(150, 700, 8_000),
                           rng.integers(700, 1500, 300)])
    cos = np.clip(0.002 * lmax - 0.1 + 0.05 * rng.standard_normal(len(lmax)), -1, 1)
    ex = 0.5 - 0.0002 * lmax + 0.05 * rng.standard_normal(len(lmax))
    return lmax, cos, ex


def test_bucket_edges_split_the_long_lengths():
    rng = np.random.default_rng(0)
    lmax, _, _ = _pairs(rng)
    up = pt.bucket_edges(lmax)
    assert np.all(np.diff(up) > 0) and up[-1] == 2**31 - 1
    q = np.quantile(lmax, np.linspace(0, 1, 9)[1:-1]).astype(np.int64)
    assert set(np.unique(q).tolist()) <= set(up.tolist())
    # the long edges above the last quantile edge and below the longest pair
    assert [e for e in up[:-1] if e > q.max()] == [e for e in pt.LONG_EDGES if q.max() < e < lmax.max()]


def test_per_bucket_thresholds_keep_every_pair_and_sparse_buckets_inherit():
    rng = np.random.default_rng(1)
    lmax, cos, ex = _pairs(rng)
    up = pt.bucket_edges(lmax)
    q = 1.0 - 1e-9  # every bucket at its extreme
    cmin = pt.per_bucket(cos, lmax, up, q, False)
    emax = pt.per_bucket(ex, lmax, up, q, True)
    lo = -1
    for k, u in enumerate(up):
        m = (lmax > lo) & (lmax <= u)
        if m.sum() >= 2000:
            assert cmin[k] == cos[m].min() and emax[k] == ex[m].max()
        else:  # too few pairs: the bucket below decides (looser than its own extreme)
            assert k > 0 and cmin[k] == cmin[k - 1] and emax[k] == emax[k - 1]
            if m.any():
                assert cmin[k] <= cos[m].min() and emax[k] >= ex[m].max()
        lo = u
    # the long buckets are tighter than one bucket over everything above 150
    tail = lmax > 150
    assert max(cmin[k] for k, u in enumerate(up) if u > 300) > cos[tail].min()


def test_calibrate_tables_fallback_buckets_skip_both_tests():
    rng = np.random.default_rng(2)
    n = 400
    L = rng.integers(80, 600, n)
    emb = rng.standard_normal((n, 8)).astype(np.float32)
    P = rng.random((n, 16)).astype(np.float32)
    i = rng.integers(0, n, 6000)
    j = rng.integers(0, n, 6000)
    keep = i != j
    i, j = i[keep], j[keep]
    upper, margins, cupper, cmin, info = pt.calibrate_tables(emb, P, L, i, j, 200, recall=0.999999,
                                                            exact_below=100)
    assert np.array_equal(upper, cupper) and upper[0] == 100
    assert margins[0] >= 1e3 and cmin[0] == -1.0
    # every calibration pair passes both tests of its bucket
    lm = np.maximum(L[i], L[j])
    b = np.searchsorted(upper, lm)
    ex = pt.radius_excess(emb, L, i, j, 200)
    cs = pt.profile_cosine(P, i, j)
    assert np.all(ex <= margins[b]) and np.all(cs >= cmin[b])
