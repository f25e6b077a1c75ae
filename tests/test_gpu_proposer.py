"""GPU tests of the learned candidate proposer: profiles and embeddings against
numpy references of the same arithmetic, training progress, and the
prefilter's epsilon lists against the exact all-pairs lists."""
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


def _kmer_ref(seqs):
    lut = np.full(256, 0, np.int64)
    for i, ch in enumerate(b"ACGT"):
        lut[ch] = i
    out = np.zeros((len(seqs), 1024), np.float64)
    for r, s in enumerate(seqs):
        c = lut[np.frombuffer(s[1:], np.uint8)]
        if len(c) < 5:
            continue
        code = np.zeros(len(c) - 4, np.int64)
        for i in range(5):
            code = code * 4 + c[i:len(c) - 4 + i]
        out[r] = np.bincount(code, minlength=1024) / len(code)
    return out


def _bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def _gelu(x):
    from scipy.special import erf
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def test_kmer_profiles_match_numpy(pkg):
    seqs = synth.evolution_tree(seed=41, n_roots=4, n_children=3, n_leaves=5) + [b"A", b"ACGTAC", b""]
    _say(f"kmer profiles n={len(seqs)}")
    with pkg.Engine(seqs) as eng:
        got = eng.kmer_profiles()
    want = _kmer_ref(seqs)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-7)


def test_embedding_matches_numpy_forward(pkg, tmp_path):
    seqs = synth.families_200bp(seed=42, n_anc=4, per_anc=30)
    _say(f"embedding n={len(seqs)}")
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=7) as enc:
        got = enc.embed(eng)
        enc.save(str(tmp_path / "enc.bin"))
    _, layers = pkg.load_encoder_weights(str(tmp_path / "enc.bin"))
    x = _bf16(_kmer_ref(seqs).astype(np.float32) * 64.0)
    (w1, b1), (w2, b2), (w3, b3) = layers
    a1 = _bf16(_gelu(x.astype(np.float64) @ _bf16(w1).T.astype(np.float64) + b1))
    a2 = _bf16(_gelu(a1.astype(np.float64) @ _bf16(w2).T.astype(np.float64) + b2))
    e = a2.astype(np.float64) @ _bf16(w3).T.astype(np.float64) + b3
    scale = np.abs(e).max()
    assert scale > 0
    np.testing.assert_allclose(got, e, rtol=0, atol=2e-2 * scale)
    # layout errors would decorrelate the embeddings entirely
    assert np.corrcoef(got.ravel(), e.ravel())[0, 1] > 0.999


def test_training_reduces_loss(pkg):
    seqs = synth.families_200bp(seed=43, n_anc=10, per_anc=40, length=160)
    rng = np.random.default_rng(0)
    n = len(seqs)
    ij = rng.integers(0, n, size=(20000, 2)).astype(np.int32)
    same = rng.integers(0, 10, size=10000) * 40
    ij[:10000, 0] = same + rng.integers(0, 40, 10000)
    ij[:10000, 1] = same + rng.integers(0, 40, 10000)
    _say(f"training on {len(ij)} pairs")
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=1) as enc:
        d = eng.pairs(ij)
        st = enc.train(eng, ij, d, epochs=8, batch=1024, lr=2e-3, eps_train=300)
        emb = enc.embed(eng)
    _say(f"loss {st['first_loss']:.4f} -> {st['final_loss']:.4f} in {st['seconds']:.2f} s")
    assert st["steps"] == 8 * 20
    assert st["final_loss"] < 0.5 * st["first_loss"]
    # within-family embedding distances are smaller than across families
    fam = np.arange(n) // 40
    a, b = ij[:, 0], ij[:, 1]
    de = np.linalg.norm(emb[a] - emb[b], axis=1)
    assert de[fam[a] == fam[b]].mean() < 0.5 * de[fam[a] != fam[b]].mean()


def test_training_is_deterministic(pkg, tmp_path):
    """Two trainings from the same seed on the same pairs give byte-identical
    weights and the same reported losses (every sum, the bias column sums
    included, runs in a fixed order), so the prefilter's candidates repeat
    from run to run."""
    seqs = synth.families_200bp(seed=43, n_anc=10, per_anc=40, length=160)
    rng = np.random.default_rng(1)
    n = len(seqs)
    ij = rng.integers(0, n, size=(6000, 2)).astype(np.int32)
    saved, losses = [], []
    with pkg.Engine(seqs) as eng:
        d = eng.pairs(ij)
        for k in range(2):
            with pkg.Encoder(seed=5) as enc:
                st = enc.train(eng, ij, d, epochs=3, batch=512, lr=2e-3, eps_train=300)
                losses.append((st["first_loss"], st["final_loss"]))
                path = tmp_path / f"enc{k}.bin"
                enc.save(str(path))
                saved.append(path.read_bytes())
    _say(f"weights {len(saved[0])} bytes, losses {losses}")
    assert saved[0] == saved[1]
    assert losses[0] == losses[1]  # per-pair losses, added in order on the host


@pytest.mark.parametrize("exact_bucket", ["1", "0"])
@pytest.mark.parametrize("tuning", [None, (3, 64, 8), (2, 16, 16)])
def test_prefilter_with_unbounded_radius_equals_exact_lists(pkg, monkeypatch, tuning, exact_bucket):
    """A huge margin proposes every pair that passes the exact length bound,
    so the verified lists must equal the exact all-pairs lists (verification
    in every kernel form and strip width, incl. the 32-column strip the short
    rows take when the context runs at 64).  With the exact bucket on, a
    margin this large switches the proposal off for every length and the
    all-pairs pass over the shortest sequences covers the whole set; off,
    every pair goes through the candidate lists."""
    monkeypatch.setenv("MNHN_PREFILTER_EXACT_BUCKET", exact_bucket)
    rng = np.random.default_rng(44)
    seqs = []
    for L0 in (70, 90, 110, 125, 140, 300, 520, 560):  # short, mid and int16-only rows
        anc = synth.random_codes(rng, L0)
        seqs += [synth.to_bytes(synth.mutate(rng, anc, 0.03, 0.01)) for _ in range(30)]
    eps = 350
    _say(f"prefilter (unbounded radius) n={len(seqs)} tuning={tuning}")
    L = np.array([len(s) for s in seqs])
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=3) as enc:
        if tuning:
            eng.set_tuning(*tuning)
        enc.margin = 1e4
        rp, col, dist, st = enc.prefilter_csr(eng, eps)
        erp, ecol, edist = eng.eps_csr(eps)
    n = len(seqs)
    iu = np.triu_indices(n, 1)
    want_cands = int((4 * np.abs(L[iu[0]] - L[iu[1]]) <= eps).sum())
    assert st["candidates"] == want_cands
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol) and np.array_equal(dist, edist)
    assert st["neighbours"] * 2 == len(col)


@pytest.mark.parametrize("split", ["2", "1", "0"])
@pytest.mark.parametrize("shape", ["lone_half_row", "mixed_counts", "long_rows"])
def test_prefilter_half_groups_equal_exact_lists(pkg, monkeypatch, split, shape):
    """Verification with target groups shared by several rows and without
    them gives the exact lists.  split 2: fp16-band and int16-band remainders
    in half pieces (two rows per group, one per half of the lanes' packed
    pairs) and quarter pieces (four rows per group); split 1: half pieces
    only.
    lone_half_row: 4 related sequences of distinct lengths hold 3, 2, 1 and 0
    candidates, so the last shared group is short of rows (its missing rows
    run no targets); mixed_counts: rows with 0-200 candidates, i.e. full
    groups, half and quarter pieces and their mixes, with query lengths that
    differ within a group (the shorter queries pre-padded) at both strip
    widths; long_rows: the same over rows longer than 512 residues, whose
    targets beyond 512 residues run the int16 form (rows near 515 residues
    hold targets of both forms), with a lone int16 half row among them."""
    monkeypatch.setenv("MNHN_VERIFY_SPLIT", split)
    monkeypatch.setenv("MNHN_PREFILTER_EXACT_BUCKET", "0")  # every pair through the candidate lists
    rng = np.random.default_rng(7)
    seqs = []
    if shape == "lone_half_row":
        anc = synth.random_codes(rng, 200)
        seqs = [synth.to_bytes(synth.mutate(rng, anc, 0.02, 0.0))[:200 - 7 * k] for k in range(4)]
    elif shape == "long_rows":
        for L0, m, ps, pi in ((515, 30, 0.03, 0.02), (600, 45, 0.01, 0.005), (700, 100, 0.005, 0.002),
                              (900, 12, 0.03, 0.02), (1300, 70, 0.004, 0.002)):
            anc = synth.random_codes(rng, L0)
            seqs += [synth.to_bytes(synth.mutate(rng, anc, ps, pi)) for _ in range(m)]
        anc = synth.random_codes(rng, 1100)
        seqs += [synth.to_bytes(synth.mutate(rng, anc, 0.02, 0.0))[:1100 - 7 * k] for k in range(4)]
    else:
        for L0, m in ((60, 40), (150, 200), (180, 70), (240, 110), (400, 35)):
            anc = synth.random_codes(rng, L0)
            seqs += [synth.to_bytes(synth.mutate(rng, anc, 0.03, 0.02)) for _ in range(m)]
    eps = 1000 if shape == "long_rows" else 300
    _say(f"prefilter half groups split={split} shape={shape} n={len(seqs)} eps={eps}")
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=3) as enc:
        enc.margin = 1e4
        rp, col, dist, st = enc.prefilter_csr(eng, eps)
        erp, ecol, edist = eng.eps_csr(eps)
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol) and np.array_equal(dist, edist)
    assert st["neighbours"] * 2 == len(col)
    if shape == "long_rows":
        assert st["cells"] > 0 and len(col) > 2 * len(seqs)


def test_prefilter_exact_bucket_lists(pkg, monkeypatch):
    """Length buckets with both tests off (radius margin >= 100, no cosine
    bound) are listed by the all-pairs pass over the shortest sequences
    instead of the candidate lists: the pairs whose longer sequence is in
    those buckets equal the exact lists' pairs, the others are a subset of
    them, and the lists equal those of the same tables run through the
    candidate lists (MNHN_PREFILTER_EXACT_BUCKET=0); the candidate count
    includes the bucket's band pairs."""
    rng = np.random.default_rng(48)
    seqs = []
    for L0, m in ((60, 50), (85, 60), (110, 40), (118, 30), (150, 40), (240, 30), (400, 20)):
        anc = synth.random_codes(rng, L0)
        seqs += [synth.to_bytes(synth.mutate(rng, anc, 0.03, 0.01)) for _ in range(m)]
    eps, cut = 300, 120
    L = np.array([len(x) for x in seqs])
    n = len(seqs)
    _say(f"prefilter exact bucket n={n} cut={cut}")
    runs = {}
    for xb in ("1", "0"):
        monkeypatch.setenv("MNHN_PREFILTER_EXACT_BUCKET", xb)
        with pkg.Engine(seqs) as eng, pkg.Encoder(seed=3) as enc:
            enc.set_margin_table(np.array([cut, 2**31 - 1], np.int32), np.array([1e3, 0.05], np.float32))
            enc.set_cosine_table(np.array([cut, 2**31 - 1], np.int32), np.array([-1.0, 0.5], np.float32))
            runs[xb] = enc.prefilter_csr(eng, eps)
            erp, ecol, edist = eng.eps_csr(eps)
    (rp, col, dist, st), (rp0, col0, dist0, st0) = runs["1"], runs["0"]
    assert np.array_equal(rp, rp0) and np.array_equal(col, col0) and np.array_equal(dist, dist0)
    assert st["candidates"] == st0["candidates"] and st["neighbours"] == st0["neighbours"]
    rows = np.repeat(np.arange(n), np.diff(rp))
    erows = np.repeat(np.arange(n), np.diff(erp))
    got = set(zip(rows.tolist(), col.tolist(), dist.tolist()))
    want = set(zip(erows.tolist(), ecol.tolist(), edist.tolist()))
    assert got <= want
    short = {t for t in want if max(L[t[0]], L[t[1]]) <= cut}
    assert short and {t for t in got if max(L[t[0]], L[t[1]]) <= cut} == short
    iu = np.triu_indices(n, 1)
    lmax = np.maximum(L[iu[0]], L[iu[1]])
    band_short = int(((4 * np.abs(L[iu[0]] - L[iu[1]]) <= eps) & (lmax <= cut)).sum())
    assert st["candidates"] >= band_short


def test_prefilter_trained_recall(pkg, oracle):
    seqs = synth.amplicon_families(seed=45, n_anc=20, per_anc=40, noise_frac=0.05)
    n = len(seqs)
    eps = 300
    rng = np.random.default_rng(2)
    ij = rng.integers(0, n, size=(30000, 2)).astype(np.int32)
    anc = np.minimum(np.arange(n) // 40, 20)
    within = rng.integers(0, 20, 15000) * 40
    ij[:15000, 0] = within + rng.integers(0, 40, 15000)
    ij[:15000, 1] = within + rng.integers(0, 40, 15000)
    _say(f"prefilter recall n={n}")
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=5) as enc:
        d = eng.pairs(ij)
        enc.train(eng, ij, d, epochs=10, batch=1024, lr=2e-3, eps_train=eps)
        enc.margin = 0.05
        rp, col, dist, st = enc.prefilter_csr(eng, eps)
        erp, ecol, edist = eng.eps_csr(eps)
    got = set(zip(np.repeat(np.arange(n), np.diff(rp)).tolist(), col.tolist()))
    want = set(zip(np.repeat(np.arange(n), np.diff(erp)).tolist(), ecol.tolist()))
    assert got <= want  # only exact neighbours are ever reported
    recall = len(got) / max(len(want), 1)
    _say(f"recall {recall:.5f}, candidates {st['candidates']}, neighbours {st['neighbours']}")
    assert recall > 0.99


def test_proposal_rule_and_margin_table_match_numpy(pkg, tmp_path):
    """Candidate count and verified lists against a numpy evaluation of the
    proposal rule (length bound, radius eps / (4 l_long) + margin(l_long) with
    a three-bucket margin table) on bf16-rounded embeddings; pairs within
    floating-point rounding of the radius are left out of the comparison."""
    seqs = synth.evolution_tree(seed=46, n_roots=8, n_children=4, n_leaves=12)
    eps = 300
    n = len(seqs)
    L = np.array([len(s) for s in seqs], dtype=np.int64)
    _say(f"proposal rule n={n}")
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=9) as enc:
        e = _bf16(enc.embed(eng)).astype(np.float64)
        d2 = ((e[:, None, :] - e[None, :, :]) ** 2).sum(-1)
        iu = np.triu_indices(n, 1)
        base = np.sqrt(np.quantile(d2[iu], [0.2, 0.35, 0.5]))
        upper = np.array([250, 400, 2**31 - 1])
        margins = base.astype(np.float32)
        enc.set_margin_table(upper, margins)
        assert abs(enc.margin - margins.max()) < 1e-7
        rp, col, dist, st = enc.prefilter_csr(eng, eps)
        erp, ecol, edist = eng.eps_csr(eps)
        enc.save(str(tmp_path / "enc.bin"))
        with pkg.Encoder(seed=1) as enc2:
            enc2.load(str(tmp_path / "enc.bin"))
            rp2, col2, dist2, st2 = enc2.prefilter_csr(eng, eps)
    tables, _ = pkg.load_encoder_weights(str(tmp_path / "enc.bin"))
    tu, tm = tables["margin"]
    assert np.array_equal(tu, upper.astype(np.int32)) and np.array_equal(tm, margins)
    assert np.all(tables["cosine"][1] < 0)
    assert st2["candidates"] == st["candidates"] and np.array_equal(col2, col)
    lmax = np.maximum(L[:, None], L[None, :])
    marg = margins[np.searchsorted(upper, lmax)]
    rad = eps / (4.0 * lmax) + marg
    nrm = (e * e).sum(1)
    tol = 1e-4 * (nrm[:, None] + nrm[None, :]) + 1e-9
    prop = (4 * np.abs(L[:, None] - L[None, :]) <= eps) & (d2 <= rad ** 2)
    border = np.abs(d2 - rad ** 2) <= tol
    prop_u, border_u = prop[iu], border[iu]
    _say(f"proposed {int(prop_u.sum())} of {len(prop_u)}, border {int(border_u.sum())}, "
         f"kernel {st['candidates']}")
    assert 0 < prop_u.sum() < len(prop_u)
    assert abs(int(st["candidates"]) - int(prop_u.sum())) <= int(border_u.sum())
    # verified lists: exactly the exact neighbours that the rule proposes
    exact = np.zeros((n, n), bool)
    exact[np.repeat(np.arange(n), np.diff(erp)), ecol] = True
    got = np.zeros((n, n), bool)
    got[np.repeat(np.arange(n), np.diff(rp)), col] = True
    want = exact & (prop | prop.T)
    sure = ~(border | border.T)
    assert np.array_equal(got & sure, want & sure)
    assert not np.any(got & ~exact)


def test_prefilter_positions_beyond_2e31(pkg, monkeypatch):
    """80,000 amplicon-like sequences with an unbounded radius propose 2.35e9
    pairs, so candidate positions pass 2^31 (wave-uniform 64-bit offsets);
    the verified lists must still equal the exact all-pairs lists.  (Every
    pair goes through the candidate lists: the exact bucket is off.)"""
    monkeypatch.setenv("MNHN_PREFILTER_EXACT_BUCKET", "0")
    full = synth.amplicon_families(seed=44)
    pick = np.sort(np.random.default_rng(44).choice(len(full), 80000, replace=False))
    seqs = [full[k] for k in pick]
    eps = 300
    _say("prefilter beyond 2^31 candidates, n=80000")
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=3) as enc:
        enc.margin = 1e4
        rp, col, dist, st = enc.prefilter_csr(eng, eps)
        _say(f"candidates {st['candidates']}, neighbours {st['neighbours']}")
        erp, ecol, edist = eng.eps_csr(eps)
    assert st["candidates"] > 2**31
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol) and np.array_equal(dist, edist)


def test_cosine_rule_matches_numpy(pkg):
    """With a cosine table set, a pair is proposed only if it also passes the
    raw 5-mer count cosine bound of its longer sequence's length bucket:
    candidate count and verified lists against numpy."""
    seqs = synth.evolution_tree(seed=47, n_roots=8, n_children=4, n_leaves=12)
    eps = 400
    n = len(seqs)
    L = np.array([len(s) for s in seqs], dtype=np.int64)
    _say(f"cosine rule n={n}")
    nk = np.maximum(L - 5, 0).astype(np.float64)
    counts = np.rint(_kmer_ref(seqs) * nk[:, None])
    nrm = np.sqrt((counts ** 2).sum(1))
    cos = (counts @ counts.T) / np.maximum(nrm[:, None] * nrm[None, :], 1e-30)
    iu = np.triu_indices(n, 1)
    upper = np.array([300, 2**31 - 1])
    cmin = np.quantile(cos[iu], [0.6, 0.8]).astype(np.float32)
    with pkg.Engine(seqs) as eng, pkg.Encoder(seed=9) as enc:
        enc.margin = 1e3  # the embedding test passes every pair
        enc.set_cosine_table(upper, cmin)
        rp, col, dist, st = enc.prefilter_csr(eng, eps)
        erp, ecol, edist = eng.eps_csr(eps)
    lmax = np.maximum(L[:, None], L[None, :])
    cm = cmin[np.searchsorted(upper, lmax)]
    prop = (4 * np.abs(L[:, None] - L[None, :]) <= eps) & (cos >= cm)
    border = np.abs(cos - cm) <= 1e-5
    _say(f"proposed {int(prop[iu].sum())} of {len(iu[0])}, border {int(border[iu].sum())}, "
         f"kernel {st['candidates']}")
    assert 0 < prop[iu].sum() < len(iu[0])
    assert abs(int(st["candidates"]) - int(prop[iu].sum())) <= int(border[iu].sum())
    exact = np.zeros((n, n), bool)
    exact[np.repeat(np.arange(n), np.diff(erp)), ecol] = True
    got = np.zeros((n, n), bool)
    got[np.repeat(np.arange(n), np.diff(rp)), col] = True
    sure = ~(border | border.T)
    assert np.array_equal(got & sure, exact & (prop | prop.T) & sure)
    assert not np.any(got & ~exact)
