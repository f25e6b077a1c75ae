"""Seeded synthetic sequence sets for the five benchmark workloads.

The reference ships a fixed-length, substitution-only simulator
(virtual_evolution.c:13-210); the workloads here need indels and uneven
lengths, so these generators are written for them.  All randomness comes from
numpy's PCG64 generator seeded per workload, so a workload is fully named by
(config, seed, size).

Mutation model (per parent site): with probability p_sub the base is replaced
by one of the three other bases (uniform); with probability p_indel there is a
single-base indel, half deletions of the site and half insertions of a uniform
random base after it.  "m% mutation with substitution:indel 4:1" means
p_sub = 0.8 m, p_indel = 0.2 m.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

ALPHABET = np.frombuffer(b"ACGT", dtype=np.uint8)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_codes(rng: np.random.Generator, length: int) -> np.ndarray:
    return rng.integers(0, 4, size=int(length), dtype=np.uint8)


def mutate(rng: np.random.Generator, parent: np.ndarray, p_sub: float,
           p_indel: float) -> np.ndarray:
    """One generation of the per-site substitution / single-base indel model."""
    n = len(parent)
    if n == 0:
        return parent.copy()
    r = rng.random(n)
    child = parent.copy()
    sub = r < p_sub
    nsub = int(sub.sum())
    if nsub:
        child[sub] = (parent[sub] + rng.integers(1, 4, size=nsub, dtype=np.uint8)) % 4
    ind = (r >= p_sub) & (r < p_sub + p_indel)
    coin = rng.random(n) < 0.5
    dels = ind & coin
    ins = ind & ~coin
    counts = 1 - dels.astype(np.int64) + ins.astype(np.int64)
    idx = np.repeat(np.arange(n), counts)
    out = child[idx]
    if len(out):
        first = np.ones(len(out), dtype=bool)
        first[1:] = idx[1:] != idx[:-1]
        nins = int((~first).sum())
        if nins:
            out[~first] = rng.integers(0, 4, size=nins, dtype=np.uint8)
    return out


def mutate_many(rng: np.random.Generator, parent: np.ndarray, p_sub, p_indel, k: int) -> List[np.ndarray]:
    """k independent children of one parent under the model of `mutate`, drawn
    as one (k, len) batch; p_sub / p_indel are scalars or one rate per child."""
    n = len(parent)
    if k <= 0:
        return []
    if n == 0:
        return [parent.copy() for _ in range(k)]
    ps = np.broadcast_to(np.asarray(p_sub, dtype=np.float64), (k,))[:, None]
    pi = np.broadcast_to(np.asarray(p_indel, dtype=np.float64), (k,))[:, None]
    r = rng.random((k, n))
    child = np.repeat(parent[None, :], k, axis=0)
    sub = r < ps
    nsub = int(sub.sum())
    if nsub:
        child[sub] = (child[sub] + rng.integers(1, 4, size=nsub, dtype=np.uint8)) % 4
    ind = (r >= ps) & (r < ps + pi)
    coin = rng.random((k, n)) < 0.5
    counts = 1 - (ind & coin).astype(np.int64) + (ind & ~coin).astype(np.int64)
    idx = np.repeat(np.arange(k * n), counts.ravel())
    out = child.ravel()[idx]
    if len(out):
        first = np.ones(len(out), dtype=bool)
        first[1:] = idx[1:] != idx[:-1]
        nins = int((~first).sum())
        if nins:
            out[~first] = rng.integers(0, 4, size=nins, dtype=np.uint8)
    return np.split(out, np.cumsum(counts.sum(1))[:-1])


def random_many(rng: np.random.Generator, lengths: Sequence[int]) -> List[np.ndarray]:
    """Uniform random sequences of the given lengths (one draw for all)."""
    lengths = np.asarray(lengths, dtype=np.int64)
    codes = rng.integers(0, 4, size=int(lengths.sum()), dtype=np.uint8)
    return np.split(codes, np.cumsum(lengths)[:-1]) if len(lengths) else []


def to_bytes(codes: np.ndarray) -> bytes:
    return ALPHABET[codes].tobytes()


def families_200bp(seed: int = 1, n_anc: int = 20, per_anc: int = 100,
            length: int = 200) -> List[bytes]:
    """Plain exact path: 20 ancestors (iid ACGT, length 200), 100 children each
    with 5% substitutions and 1% single-base indels; N = 2,000."""
    rng = _rng(seed)
    out = []
    for _ in range(n_anc):
        anc = random_codes(rng, length)
        for _ in range(per_anc):
            out.append(to_bytes(mutate(rng, anc, 0.05, 0.01)))
    return out


def evolution_tree(seed: int = 2, n_roots: int = 100, n_children: int = 10,
            n_leaves: int = 20, median: float = 300.0, sigma: float = 0.4,
            lo: int = 100, hi: int = 600) -> List[bytes]:
    """Uneven-length batching: 3-level tree.  Roots have lognormal lengths
    (median 300, sigma 0.4, clipped to [100, 600]); 10 children per root at 8%
    mutation, 20 leaves per child at 3% mutation (sub:indel 4:1).  100 roots
    give N = 20,000 leaves."""
    rng = _rng(seed)
    out = []
    for _ in range(n_roots):
        L = int(np.clip(np.rint(rng.lognormal(np.log(median), sigma)), lo, hi))
        root = random_codes(rng, L)
        for _ in range(n_children):
            child = mutate(rng, root, 0.8 * 0.08, 0.2 * 0.08)
            for _ in range(n_leaves):
                out.append(to_bytes(mutate(rng, child, 0.8 * 0.03, 0.2 * 0.03)))
    return out


def amplicon_families(seed: int = 3, n_anc: int = 500, per_anc: int = 200,
            noise_frac: float = 0.05, lo: int = 150, hi: int = 300) -> List[bytes]:
    """Amplicon-scale: 500 ancestors with length uniform in [150, 300]; 200
    descendants each at a mutation rate drawn uniformly in [2%, 10%]
    (sub:indel 4:1); then 5% extra uniform-random sequences of the same length
    range.  N = 100,000 + 5,000 = 105,000."""
    rng = _rng(seed)
    out = []
    for _ in range(n_anc):
        anc = random_codes(rng, rng.integers(lo, hi + 1))
        m = rng.uniform(0.02, 0.10, size=per_anc)
        out += [to_bytes(c) for c in mutate_many(rng, anc, 0.8 * m, 0.2 * m, per_anc)]
    n_noise = int(round(noise_frac * n_anc * per_anc))
    out += [to_bytes(c) for c in random_many(rng, rng.integers(lo, hi + 1, size=n_noise))]
    return out


def pareto_lengths(rng: np.random.Generator, n: int, alpha: float = 1.5, xm: float = 80.0,
                   hi: int = 1500) -> np.ndarray:
    """Pareto(alpha, x_m) lengths clipped to [x_m, hi]."""
    return np.minimum(np.rint(xm * (1.0 + rng.pareto(alpha, size=n))), hi).astype(np.int64)


# tree shape of heavy_tail_mix: roots x children x grandchildren = 2,000 leaf clusters
MIX_ROOTS, MIX_PER_ROOT, MIX_PER_CHILD = 250, 4, 2


def heavy_tail_mix(seed: int = 5, n: int = 500_000, n_roots: int = MIX_ROOTS, per_root: int = MIX_PER_ROOT,
            per_child: int = MIX_PER_CHILD, tree_frac: float = 0.70, repeat_frac: float = 0.20,
            n_monomer_families: int = 200, monomer: int = 171, with_groups: bool = False):
    """End-to-end learned-prefilter workload (heavy-tailed lengths 80-1500):
      70% from a 4-level tree with 2,000 leaf clusters: 250 roots (Pareto
         lengths, alpha 1.5, x_m 80, clipped to 1,500; each root is its own
         length draw) -> 4 children at 8% -> 2 grandchildren at 4% (= 2,000
         leaf-cluster ancestors) -> members at 2%;
      20% tandem repeats: 2-8 copies of one of 200 random 171-bp monomers, every
         copy at 3% divergence;
      10% uniform random sequences with Pareto lengths.
    (mutation: sub:indel 4:1).  with_groups=True also returns one label per
    sequence: its leaf cluster, its monomer family, or a unique label for noise.
    Leaf clusters are numbered root-major: leaf g descends from child
    g // per_child and root g // (per_root * per_child)."""
    rng = _rng(seed)
    n_tree = int(round(tree_frac * n))
    n_rep = int(round(repeat_frac * n))
    n_noise = n - n_tree - n_rep
    leaves = []
    for L in pareto_lengths(rng, n_roots):
        root = random_codes(rng, L)
        for child in mutate_many(rng, root, 0.8 * 0.08, 0.2 * 0.08, per_root):
            leaves += mutate_many(rng, child, 0.8 * 0.04, 0.2 * 0.04, per_child)
    # sequence k of the tree part descends from leaf k % 2,000
    out: List[bytes] = [b""] * n
    nl = len(leaves)
    for j, leaf in enumerate(leaves):
        members = range(j, n_tree, nl)
        for k, c in zip(members, mutate_many(rng, leaf, 0.8 * 0.02, 0.2 * 0.02, len(members))):
            out[k] = to_bytes(c)
    # sequence n_tree + k repeats monomer k % 200, 2-8 copies at 3% each
    monomers = [random_codes(rng, monomer) for _ in range(n_monomer_families)]
    copies = rng.integers(2, 9, size=n_rep)
    for f, mono in enumerate(monomers):
        members = np.arange(f, n_rep, n_monomer_families)
        parts = mutate_many(rng, mono, 0.8 * 0.03, 0.2 * 0.03, int(copies[members].sum()))
        at = 0
        for k in members:
            out[n_tree + k] = to_bytes(np.concatenate(parts[at:at + copies[k]]))
            at += copies[k]
    for k, c in enumerate(random_many(rng, pareto_lengths(rng, n_noise))):
        out[n_tree + n_rep + k] = to_bytes(c)
    if not with_groups:
        return out
    k = np.arange(n)
    groups = np.where(k < n_tree, k % len(leaves),
                      np.where(k < n_tree + n_rep, len(leaves) + (k - n_tree) % n_monomer_families,
                               len(leaves) + n_monomer_families + k))
    return out, groups


def amplicon_family_labels(n_anc: int = 500, per_anc: int = 200, noise_frac: float = 0.05) -> np.ndarray:
    """Ancestor label of every sequence of amplicon_families() with the same sizes
    (noise sequences get unique labels)."""
    n_desc = n_anc * per_anc
    n = n_desc + int(round(noise_frac * n_desc))
    k = np.arange(n)
    return np.where(k < n_desc, k // per_anc, n_anc + k)


def labelled_pairs(rng: np.random.Generator, n: int, group: np.ndarray, m: int,
                   within_frac: float = 0.5) -> np.ndarray:
    """m index pairs: within_frac of them inside the same group label, the
    rest uniform random (no self pairs)."""
    m_in = int(round(within_frac * m))
    order = np.argsort(group, kind="stable")
    g_sorted = group[order]
    starts = np.searchsorted(g_sorted, g_sorted, side="left")
    ends = np.searchsorted(g_sorted, g_sorted, side="right")
    a = rng.integers(0, n, size=m_in)
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    pa = inv[a]
    lo, hi = starts[pa], ends[pa]
    b = order[lo + (rng.random(m_in) * (hi - lo)).astype(np.int64)]
    rnd = rng.integers(0, n, size=(m - m_in, 2))
    ij = np.concatenate([np.stack([a, b], 1), rnd]).astype(np.int32)
    ij = ij[ij[:, 0] != ij[:, 1]]
    return ij


def write_fasta(path: str, seqs: Sequence[bytes], width: int = 50) -> None:
    """'>sequence_<i>' headers with lines wrapped at `width` columns."""
    with open(path, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b">sequence_%d\n" % i)
            for k in range(0, len(s), width):
                f.write(s[k:k + width] + b"\n")
