"""GPU parity tests of the k-mer counter (mnhn_kmer.h) and the fasta2kmer
drop-in: tables bit-identical to the oracle's (whose counting rules are pinned
to the reference's fasta2kmer fixtures), the CLI byte-identical to the
reference on every fixture, and properties at larger sizes."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, golden_arg, vector_cases

pytestmark = pytest.mark.gpu

KMER_CASES = vector_cases({"fasta2kmer"})


def _random_seqs(rng, n, lo, hi, alphabet):
    a = np.frombuffer(alphabet, np.uint8)
    return [bytes(rng.choice(a, size=int(rng.integers(lo, hi))).astype(np.uint8)) for _ in range(n)]


@pytest.mark.parametrize("k,protein", [(1, False), (2, False), (3, False), (5, False), (7, False), (8, False),
                                       (1, True), (2, True), (3, True), (4, True)])
def test_kmer_tables_match_oracle(pkg, oracle, k, protein):
    """Both device paths (LDS histogram up to 32768 columns, device-memory
    rows above: DNA k = 8, protein k = 4), foreign bytes and short
    sequences."""
    rng = np.random.default_rng(k * 10 + protein)
    alpha = b"ACDEFGHIKLMNPQRSTVWYXB*" if protein else b"ACGTN"
    n = 40 if (protein and k == 4) else 120
    seqs = _random_seqs(rng, n, 0, 700, alpha) + [b"", b"A" * k, b"A" * (k + 1)]
    print(f"kmer table k={k} protein={protein} n={len(seqs)}", flush=True)
    got = pkg.kmer_frequencies(seqs, k, protein)
    assert np.array_equal(got, oracle.kmer_frequencies(seqs, k, protein))


def test_kmer_counts_wrap(pkg, oracle):
    """A column with more than 32767 windows wraps like the reference's short
    counter; a long sequence spans many strides of the workgroup."""
    seqs = [b"A" * 70000, b"AC" * 40000, b"ACGT" * 20000 + b"N" + b"T" * 33000]
    for k in (1, 3, 8):
        got = pkg.kmer_frequencies(seqs, k, False)
        assert np.array_equal(got, oracle.kmer_frequencies(seqs, k, False)), k
    assert pkg.kmer_frequencies(seqs[:1], 1, False)[0, 0] == np.int16(69999 - 65536)


def test_kmer_properties_at_size(pkg):
    """50,000 sequences of 100-2,000 bases, k = 5: every row sums to the
    number of ACGT windows below len - k, and 200 sampled rows equal a numpy
    count."""
    rng = np.random.default_rng(4)
    n = 50000
    lens = rng.integers(100, 2000, n)
    codes = rng.integers(0, 4, int(lens.sum())).astype(np.uint8)
    codes[rng.random(len(codes)) < 0.001] = 4
    letters = np.frombuffer(b"ACGTN", np.uint8)[codes]
    offs = np.concatenate([[0], np.cumsum(lens)])
    seqs = [letters[offs[i]:offs[i + 1]].tobytes() for i in range(n)]
    print("kmer properties n=50000 k=5", flush=True)
    got = pkg.kmer_frequencies(seqs, 5, False)
    assert got.shape == (n, 1024)
    for i in rng.choice(n, 200, replace=False):
        c = codes[offs[i]:offs[i + 1]].astype(np.int64)
        w = len(c) - 5
        ok = np.ones(max(w, 0), bool)
        col = np.zeros(max(w, 0), np.int64)
        for p in range(5):
            ok &= c[p:p + w] < 4
            col |= (c[p:p + w] & 3) << (2 * p)
        want = np.bincount(col[ok], minlength=1024)
        assert np.array_equal(got[i].astype(np.int64), want), i
    assert int(got.astype(np.int64).sum()) > 0


@pytest.mark.parametrize("name,meta", KMER_CASES, ids=[c[0] for c in KMER_CASES])
def test_fasta2kmer_cli_matches_reference(name, meta, tmp_path):
    exe = os.path.join(PKG_DIR, "bin", "fasta2kmer")
    assert os.path.exists(exe), f"{exe} not built"
    a = meta["argv"]
    src = golden_arg(a[0])
    dst = tmp_path / "in.fasta"
    shutil.copyfile(src, dst)
    print(f"fasta2kmer case {name}", flush=True)
    r = subprocess.run([exe, str(dst)] + a[1:], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == meta["returncode"], r.stderr.decode()[-2000:]
    assert r.stdout == open(os.path.join(GOLDEN, "vectors", name, "stdout.txt"), "rb").read()
