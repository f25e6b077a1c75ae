"""Training and calibration of the learned candidate proposer.

The encoder (5-mer counts -> 128-d, bf16 MFMA kernels in csrc/proposer.hip) is
trained on pairs labelled with exact Smith-Waterman distances from the
engine itself; its two proposal tests (embedding radius, count cosine) get one
threshold per length bucket, calibrated on every exact neighbour pair among
held-out sequences the encoder never saw.  Proposals only choose which pairs
the exact kernel scores: the lists hold exact distances only, so a pair that
is not proposed is missing, never wrong.  Also here: the recall / partition
checks of prefilter lists against exact lists.
"""
import time

import numpy as np

import synth


def _bf16_round(x):
    """float32 -> nearest-even bf16, returned as float32 (the embeddings the
    pairwise kernel reads)."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    return ((a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def radius_excess(emb16, L, i, j, eps, match=4, chunk=1 << 20):
    """||e_i - e_j|| - eps / (match max(l_i, l_j)) per pair: a neighbour pair
    is proposed iff its excess is <= the encoder's margin."""
    out = np.empty(len(i), np.float32)
    for s in range(0, len(i), chunk):
        a, b = i[s:s + chunk], j[s:s + chunk]
        de = np.sqrt(((emb16[a] - emb16[b]) ** 2).sum(1))
        out[s:s + chunk] = de - eps / (match * np.maximum(np.maximum(L[a], L[b]), 1))
    return out


def sample_training_pairs(rng, group_levels, n, m):
    """m labelled-pair indices: for each (groups, fraction) level that fraction
    drawn inside the same group, the rest uniform random."""
    parts, used = [], 0
    for groups, frac in group_levels:
        k = int(round(frac * m))
        parts.append(synth.labelled_pairs(rng, n, groups, k, within_frac=1.0))
        used += k
    parts.append(synth.labelled_pairs(rng, n, np.arange(n), m - used, within_frac=0.0))
    ij = np.concatenate(parts)
    return ij[rng.permutation(len(ij))]


def profile_cosine(P, i, j, chunk=1 << 18):
    """Cosine of the 5-mer profiles of pairs (i, j) (rows of P)."""
    out = np.empty(len(i), np.float32)
    nrm = np.sqrt((P.astype(np.float64) ** 2).sum(1))
    for s in range(0, len(i), chunk):
        a, b = i[s:s + chunk], j[s:s + chunk]
        out[s:s + chunk] = (P[a] * P[b]).sum(1) / np.maximum(nrm[a] * nrm[b], 1e-30)
    return out


# bucket edges beyond the quantile edges: neighbour pairs crowd at the short
# lengths, so the quantile buckets alone leave every longer pair in one bucket
# whose thresholds are set by its shortest pairs
LONG_EDGES = (150, 200, 300, 450, 700, 1000)


def bucket_edges(lmax, buckets=8, long_edges=LONG_EDGES):
    """Upper edges of the length buckets (longer sequence of a pair): the
    quantiles of lmax in `buckets` equal parts, then long_edges above the last
    quantile and below the longest pair."""
    qs = np.unique(np.quantile(lmax, np.linspace(0, 1, buckets + 1)[1:-1]).astype(np.int64))
    top = qs[-1] if len(qs) else 0
    extra = [e for e in long_edges if top < e < lmax.max()]
    return np.concatenate([qs, extra, [2**31 - 1]]).astype(np.int64)


def per_bucket(values, lmax, upper, q, high, min_pairs=2000):
    """Per length bucket: the q-quantile of values (high=True) or the
    (1-q)-quantile (high=False), the extreme when the quantile rests on fewer
    than 10 pairs.  A bucket of fewer than min_pairs pairs takes the value of
    the bucket below it (the looser one: neighbour pairs of longer sequences
    keep more of their k-mers and sit closer in the embedding), the first
    bucket the extreme over all pairs."""
    out, lo = [], -1
    for u in upper:
        v = values[(lmax > lo) & (lmax <= u)]
        if len(v) < min_pairs:
            out.append(out[-1] if out else float(values.max() if high else values.min()))
        elif len(v) * (1 - q) < 10:
            out.append(float(v.max() if high else v.min()))
        else:
            out.append(float(np.quantile(v, q if high else 1 - q)))
        lo = u
    return out


def calibrate_tables(emb16, P, L, i, j, eps, recall=0.9999, slack=0.005, exact_below=0):
    """Margin and cosine tables from neighbour pairs (i, j) of held-out
    sequences, per length bucket of the longer sequence (buckets of equal pair
    counts).  Each test keeps all but (1 - recall) / 2 of the pairs of its
    bucket, so the pair passes both with probability >= recall.  Pairs whose
    longer sequence is at most exact_below long get their own buckets with
    both tests off: every pair of them that the length bound allows goes to
    the exact kernel."""
    q = 1 - (1 - recall) / 2
    if len(i) == 0:
        one = np.array([2**31 - 1], np.int32)
        return one, np.array([slack], np.float32), one, np.array([-1.0], np.float32), {}
    ex = radius_excess(emb16, L, i, j, eps)
    lmax = np.maximum(L[i], L[j])
    upper = bucket_edges(lmax)
    if exact_below > 0:
        upper = np.unique(np.concatenate([[exact_below], upper[upper > exact_below]])).astype(np.int64)
    fallback = upper <= exact_below
    margins = [1e3 if fb else max(m, 0.0) + slack
               for m, fb in zip(per_bucket(ex, lmax, upper, q, True), fallback)]
    info = {"pairs": int(len(ex)), "excess_q50": float(np.quantile(ex, 0.5)),
            "excess_q999": float(np.quantile(ex, 0.999)), "excess_max": float(ex.max()),
            "max_len": upper.tolist(), "margins": margins}
    if P is None:
        cmin = [-1.0] * len(upper)
    else:
        cs = profile_cosine(P, i, j)
        cmin = [-1.0 if fb else max(c - slack, -1.0)
                for c, fb in zip(per_bucket(cs, lmax, upper, q, False), fallback)]
        info.update({"cos_q001": float(np.quantile(cs, 0.001)), "cos_min": float(cs.min()),
                     "cos_table": cmin})
    return (upper.astype(np.int32), np.array(margins, np.float32), upper.astype(np.int32),
            np.array(cmin, np.float32), info)


def train_encoder(pkg, seqs, group_levels, holdout_groups, n_pairs, eps, seed, epochs, log,
                  use_cosine=True, recall=0.9999, exact_below=0):
    """Label n_pairs pairs of seqs with the exact kernel and train an encoder on
    the pairs between sequences outside a held-out 10% of `holdout_groups`.
    The margin and cosine tables come from every neighbour pair (d <= eps,
    exact all-pairs lists) among the held-out sequences, which the encoder
    never saw."""
    rng = np.random.default_rng(seed)
    n = len(seqs)
    L = np.array([len(x) for x in seqs], dtype=np.int64)
    ug = np.unique(holdout_groups)
    held = np.isin(holdout_groups, rng.choice(ug, max(1, len(ug) // 10), replace=False))
    ij = sample_training_pairs(rng, group_levels, n, n_pairs)
    eng = pkg.Engine(seqs)
    t0 = time.perf_counter()
    d = eng.pairs(ij)
    label_s = time.perf_counter() - t0
    tr = ~held[ij[:, 0]] & ~held[ij[:, 1]]
    cal = held[ij[:, 0]] & held[ij[:, 1]]
    log(f"labelled {len(ij)} pairs in {label_s:.2f} s; training pairs {int(tr.sum())}, "
        f"held-out pairs {int(cal.sum())}; neighbours (d <= {eps}): {int((d <= eps).sum())}")
    enc = pkg.Encoder(seed=seed)
    st = enc.train(eng, ij[tr], d[tr], epochs=epochs, batch=1024, lr=2e-3, eps_train=eps,
                   neighbor_weight=4.0, seed=seed, weight_decay=0.05, lr_decay=True,
                   far_weight=1.0, far_gap=0.3)
    log(f"trained: loss {st['first_loss']:.4f} -> {st['final_loss']:.4f}, {st['steps']} steps, "
        f"{st['seconds']:.1f} s")
    emb16 = _bf16_round(enc.embed(eng))
    P = eng.kmer_profiles() if use_cosine else None
    # calibration set: every neighbour pair (exact all-pairs lists) among the
    # held-out sequences, at most 60,000 of them
    hidx = np.nonzero(held)[0]
    if len(hidx) > 60000:
        hidx = np.sort(rng.choice(hidx, 60000, replace=False))
    with pkg.Engine([seqs[k] for k in hidx]) as heng:
        hrp, hcol, _ = heng.eps_csr(eps)
    hrows = np.repeat(np.arange(len(hidx)), np.diff(hrp))
    up = hrows < hcol
    ci, cj = hidx[hrows[up]], hidx[hcol[up]]
    del hrp, hcol, hrows, up
    upper, margins, cupper, cmin, cinfo = calibrate_tables(emb16, P, L, ci, cj, eps, recall=recall,
                                                           exact_below=exact_below)
    del P
    enc.set_margin_table(upper, margins)
    enc.set_cosine_table(cupper, cmin)
    log(f"tables from {len(ci)} neighbour pairs of {len(hidx)} held-out sequences: {cinfo}")
    eng.close()
    info = {"labelled_pairs": int(len(ij)), "train_pairs": int(tr.sum()), "label_s": label_s,
            "train_s": st["seconds"], "epochs": epochs, "steps": st["steps"],
            "loss_first": st["first_loss"], "loss_final": st["final_loss"],
            "calibration": cinfo, "calibration_sequences": int(len(hidx))}
    return enc, info


def same_partition(a, b):
    return sorted(tuple(sorted(c)) for c in a) == sorted(tuple(sorted(c)) for c in b)


def pair_set_recall(rp_p, col_p, dist_p, rp_e, col_e, dist_e, sample_rows=20000, seed=0):
    """Recall of the prefilter lists against the exact lists, whether every
    prefilter entry (row, col, d) is an exact entry, and the exact entries the
    prefilter missed, checked on a random sample of rows (all rows when there
    are fewer).  Columns ascend within each row in both lists."""
    n = len(rp_e) - 1
    rows = np.arange(n) if n <= sample_rows else \
        np.random.default_rng(seed).choice(n, sample_rows, replace=False)
    subset = True
    missed = []
    for r in rows:
        a, b = slice(rp_p[r], rp_p[r + 1]), slice(rp_e[r], rp_e[r + 1])
        ce, de, cp = col_e[b], dist_e[b], col_p[a]
        pos = np.searchsorted(ce, cp)
        ok = (pos < len(ce))
        ok &= ce[np.minimum(pos, len(ce) - 1)] == cp if len(ce) else False
        if not np.all(ok) or not np.array_equal(de[pos], dist_p[a]):
            subset = False
        if len(cp) < len(ce):
            for c in np.setdiff1d(ce, cp):
                missed.append((int(r), int(c)))
    return len(col_p) / max(len(col_e), 1), subset, missed


def lists_subset_and_recall(rp_p, col_p, dist_p, rp_e, col_e, dist_e):
    """Every prefilter entry (row, col, d) checked against the exact lists
    (all rows, vectorised): (every entry is an exact entry with the same d,
    recall = |prefilter| / |exact|).  Columns ascend within rows in both."""
    n = len(rp_e) - 1
    re = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp_e))
    rpf = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp_p))
    ke = re * n + np.asarray(col_e, np.int64)
    kp = rpf * n + np.asarray(col_p, np.int64)
    pos = np.searchsorted(ke, kp)
    ok = pos < len(ke)
    ok[ok] &= ke[pos[ok]] == kp[ok]
    subset = bool(ok.all()) and bool(np.array_equal(np.asarray(dist_e)[pos[ok]], np.asarray(dist_p)[ok]))
    return subset, len(col_p) / max(len(col_e), 1)


def proposer100k_workload(pkg, log, eps=300, pairs=2_000_000, epochs=3, train_anc=20000, train_per=20,
                          use_cosine=True, recall=0.9999, exact_below=0, n=100_000):
    """The 100k held-out proposer workload: an encoder trained on `pairs`
    exact-labelled pairs (50% within-ancestor, 50% random) of the
    amplicon-scale generator (lengths 150-300, 2-10% mutation, 5% noise) drawn
    with train_anc ancestors x train_per (seed 4), and n sequences of another
    draw (seed 44) to evaluate on.  Returns (encoder, training info,
    evaluation sequences, description)."""
    train = synth.amplicon_families(seed=4, n_anc=train_anc, per_anc=train_per)
    groups = synth.amplicon_family_labels(n_anc=train_anc, per_anc=train_per)
    enc, info = train_encoder(pkg, train, [(groups, 0.5)], groups, pairs, eps, 4, epochs, log,
                              use_cosine=use_cosine, recall=recall, exact_below=exact_below)
    del train
    full = synth.amplicon_families(seed=44)
    pick = np.sort(np.random.default_rng(44).choice(len(full), n, replace=False))
    seqs = [full[k] for k in pick]
    workload = (f"learned proposer: encoder trained on {pairs} exact-labelled pairs "
                f"(50% within-ancestor, 50% random) of the amplicon-scale generator drawn "
                f"with {train_anc} ancestors x {train_per} (seed 4), evaluated on "
                f"{len(seqs)} sequences of another draw (seed 44)")
    return enc, info, seqs, workload


def mix500k_workload(pkg, log, eps=200, pairs=2_000_000, epochs=3, train_draws=4, train_n=100_000,
                     use_cosine=True, recall=0.999999, exact_below=100, n=500_000):
    """The heavy-tailed end-to-end workload: an encoder trained on `pairs`
    exact-labelled pairs of train_draws other draws (seeds 50, 51, ...) of
    train_n sequences each (40% same leaf cluster or monomer family, 15% same
    child, 10% same root, 35% random), and n sequences of the seed-5 draw to
    evaluate on.  Pairs whose longer sequence is at most exact_below long skip
    the proposal tests (every such pair the length bound allows is verified).
    Returns (encoder, training info, evaluation sequences, description)."""
    parts, labels = [], []
    for k in range(train_draws):
        tr, g = synth.heavy_tail_mix(seed=50 + k, n=train_n, with_groups=True)
        parts += tr
        labels.append((g, k))
    big = 1 << 40
    leaf = np.concatenate([g + k * big for g, k in labels])
    n_leaf = synth.MIX_ROOTS * synth.MIX_PER_ROOT * synth.MIX_PER_CHILD
    child = np.concatenate([np.where(g < n_leaf, g // synth.MIX_PER_CHILD, g) + k * big for g, k in labels])
    root = np.concatenate([np.where(g < n_leaf, g // (synth.MIX_PER_ROOT * synth.MIX_PER_CHILD), g) + k * big
                           for g, k in labels])
    enc, info = train_encoder(pkg, parts, [(leaf, 0.4), (child, 0.15), (root, 0.1)], root, pairs, eps, 5,
                              epochs, log, use_cosine=use_cosine, recall=recall, exact_below=exact_below)
    del parts
    seqs = synth.heavy_tail_mix(seed=5, n=n)
    workload = ("end-to-end learned-prefilter DBSCAN: heavy-tailed lengths 80-1500 "
                "(Pareto 1.5), 70% 4-level tree with 2,000 leaf clusters, 20% tandem "
                "repeats of 171-bp monomers (2-8 copies, 3%), 10% noise; encoder trained "
                f"on {train_draws} other draws of {train_n}")
    return enc, info, seqs, workload
