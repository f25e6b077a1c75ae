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


def to_bytes(codes: np.ndarray) -> bytes:
    return ALPHABET[codes].tobytes()


def config1(seed: int = 1, n_anc: int = 20, per_anc: int = 100,
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


def config2(seed: int = 2, n_roots: int = 100, n_children: int = 10,
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


def config3(seed: int = 3, n_anc: int = 500, per_anc: int = 200,
            noise_frac: float = 0.05, lo: int = 150, hi: int = 300) -> List[bytes]:
    """Amplicon-scale: 500 ancestors with length uniform in [150, 300]; 200
    descendants each at a mutation rate drawn uniformly in [2%, 10%]
    (sub:indel 4:1); then 5% extra uniform-random sequences of the same length
    range.  N = 100,000 + 5,000 = 105,000."""
    rng = _rng(seed)
    out = []
    for _ in range(n_anc):
        anc = random_codes(rng, rng.integers(lo, hi + 1))
        for _ in range(per_anc):
            m = rng.uniform(0.02, 0.10)
            out.append(to_bytes(mutate(rng, anc, 0.8 * m, 0.2 * m)))
    n_noise = int(round(noise_frac * n_anc * per_anc))
    for _ in range(n_noise):
        out.append(to_bytes(random_codes(rng, rng.integers(lo, hi + 1))))
    return out


def write_fasta(path: str, seqs: Sequence[bytes], width: int = 50) -> None:
    """'>sequence_<i>' headers with lines wrapped at `width` columns."""
    with open(path, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b">sequence_%d\n" % i)
            for k in range(0, len(s), width):
                f.write(s[k:k + width] + b"\n")
