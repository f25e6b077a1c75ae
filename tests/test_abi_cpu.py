"""CPU tests of the C ABI library and its host code (no GPU calls):
symbols, FASTA parsing, split-set and cluster-file formats, relations."""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            prev = text[:m.start()].rstrip()
            if name in ("if", "sizeof", "defined") or prev.endswith("typedef"):
                continue
            if re.search(r"[\w\*]\s*$", prev) and not prev.endswith("#define"):
                names.add(name)
    names.discard("dbscanner")  # function-pointer parameter name
    return names


def test_headers_declare_the_exported_list(pkg):
    assert _declared_functions() == (set(pkg.ABI_SYMBOLS) | set(pkg.HOST_SYMBOLS)
                                     | set(pkg.PROPOSER_SYMBOLS) | set(pkg.VEC_SYMBOLS))


def test_library_loads_and_exports_every_symbol(pkg):
    lib = pkg.load_library()
    for name in pkg.ABI_SYMBOLS + pkg.HOST_SYMBOLS + pkg.PROPOSER_SYMBOLS + pkg.VEC_SYMBOLS:
        assert hasattr(lib, name), name
    assert lib.mnhn_sw_strerror(0) == b"ok"
    assert lib.mnhn_sw_strerror(5).startswith(b"epsilon above")


def test_create_rejects_bad_arguments(pkg):
    lib = pkg.load_library()
    out = ctypes.c_void_p()
    assert lib.mnhn_sw_create(None, None, -1, None, ctypes.byref(out)) == 1
    assert lib.mnhn_sw_create(None, None, 0, None, None) == 1


# ---- host code through ctypes ------------------------------------------------

class Dataset(ctypes.Structure):
    _fields_ = [("sequences", ctypes.POINTER(ctypes.c_char_p)),
                ("sequence_lengths", ctypes.POINTER(ctypes.c_size_t)),
                ("binary_sequences", ctypes.c_void_p), ("n_values", ctypes.c_int),
                ("values", ctypes.c_void_p), ("n_dimensions", ctypes.c_int),
                ("max_sequence_length", ctypes.c_size_t)]


class Cluster(ctypes.Structure):
    _fields_ = [("members", ctypes.POINTER(ctypes.c_int)), ("n_members", ctypes.c_int),
                ("id", ctypes.c_int)]


class SplitSet(ctypes.Structure):
    _fields_ = [("n_clusters", ctypes.c_int), ("clusters", ctypes.POINTER(Cluster))]


class Conn(ctypes.Structure):
    _fields_ = [("connections", ctypes.POINTER(ctypes.c_int)), ("n_connections", ctypes.c_int)]


@pytest.fixture(scope="module")
def host(pkg):
    lib = pkg.load_library()
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    lib.dataset_from_fasta.restype = Dataset
    lib.dataset_from_fasta.argtypes = [ctypes.c_void_p]
    lib.free_sequences_from_dataset.argtypes = [Dataset]
    lib.store_split_set.argtypes = [ctypes.c_char_p, SplitSet]
    lib.read_split_set.restype = SplitSet
    lib.read_split_set.argtypes = [ctypes.c_char_p]
    lib.create_cluster_files.argtypes = [ctypes.c_char_p, SplitSet, Dataset]
    lib.generate_split_set_relation.restype = ctypes.POINTER(Conn)
    lib.generate_split_set_relation.argtypes = [SplitSet, SplitSet]
    return lib, libc


def _read_c(host, path):
    lib, libc = host
    f = libc.fopen(path.encode(), b"r")
    ds = lib.dataset_from_fasta(f)
    libc.fclose(f)
    seqs = [ds.sequences[i][:ds.sequence_lengths[i]] for i in range(ds.n_values)]
    assert ds.max_sequence_length == max([len(s) for s in seqs], default=0)
    lib.free_sequences_from_dataset(ds)
    return seqs


def test_fasta_reader_matches_oracle_on_inputs(host, oracle):
    d = os.path.join(GOLDEN, "inputs")
    for fn in sorted(os.listdir(d)):
        assert _read_c(host, os.path.join(d, fn)) == oracle.read_fasta(os.path.join(d, fn)), fn


def test_fasta_reader_edge_cases(host, oracle, tmp_path):
    p = tmp_path / "edge.fasta"
    p.write_bytes(b">a\nacgt  xx\n\n  tt\n>b\n>c\n\t\nGgNn-\n>d no newline at end\nAC")
    got = _read_c(host, str(p))
    assert got == oracle.read_fasta(str(p))
    # blank / whitespace-only lines repeat the previous token (sscanf quirk)
    assert got == [b"ACGTACGTTT", b"", b"TTGGNN-", b"AC"]


def _mk_split(clusters):
    arr = (Cluster * max(len(clusters), 1))()
    keep = []
    for k, (members, cid) in enumerate(clusters):
        m = (ctypes.c_int * max(len(members), 1))(*members)
        keep.append(m)
        arr[k] = Cluster(ctypes.cast(m, ctypes.POINTER(ctypes.c_int)), len(members), cid)
    s = SplitSet(len(clusters), ctypes.cast(arr, ctypes.POINTER(Cluster)))
    return s, (arr, keep)


def test_split_set_bytes_and_roundtrip(host, oracle, tmp_path):
    lib, _ = host
    cl = [([3, 1, 2], 0), ([], 1), (list(range(40, 10, -1)), 7)]
    s, keep = _mk_split(cl)
    path = str(tmp_path / "ss0000").encode()
    lib.store_split_set(path, s)
    data = open(path, "rb").read()
    want = oracle.split_set_bytes(oracle.SplitSet([oracle.Cluster(m, i) for m, i in cl]))
    assert data == want
    assert data[:9] == b"SPLIT_SET" and struct.unpack("<i", data[9:13])[0] == 3
    r = lib.read_split_set(path)
    assert r.n_clusters == 3
    assert [r.clusters[k].members[i] for k in range(3) for i in range(r.clusters[k].n_members)] == \
        [x for m, _ in cl for x in m]
    # no O_TRUNC: a longer pre-existing file keeps its tail
    longer = tmp_path / "ss0001"
    longer.write_bytes(b"Z" * (len(want) + 5))
    lib.store_split_set(str(longer).encode(), s)
    assert longer.read_bytes() == want + b"Z" * 5


def test_cluster_files_match_oracle(host, oracle, tmp_path):
    lib, libc = host
    path = os.path.join(GOLDEN, "inputs", "c2_small.fasta")
    f = libc.fopen(path.encode(), b"r")
    ds = lib.dataset_from_fasta(f)
    libc.fclose(f)
    seqs = oracle.read_fasta(path)
    cl = [([5, 0, 17, 100], 0), ([250, 3], 1)]
    s, keep = _mk_split(cl)
    prefix = str(tmp_path / "out").encode()
    lib.create_cluster_files(prefix, s, ds)
    for k, (m, i) in enumerate(cl):
        got = open(str(tmp_path / ("out-%03d" % k)), "rb").read()
        assert got == oracle.cluster_file_text(oracle.Cluster(m, i), seqs)
    lib.free_sequences_from_dataset(ds)


def test_relation_and_sort_side_effect(host, oracle):
    lib, _ = host
    rng = np.random.default_rng(1)
    old = [rng.permutation(100)[:k].tolist() for k in (30, 15, 25)]
    new = [old[0] + old[2][:10], old[1][:5]]
    so, ko = _mk_split([(m, i) for i, m in enumerate(old)])
    sn, kn = _mk_split([(m, i) for i, m in enumerate(new)])
    conn = lib.generate_split_set_relation(so, sn)
    got = [[conn[i].connections[k] for k in range(conn[i].n_connections)] for i in range(2)]
    O = [oracle.Cluster(list(m), i) for i, m in enumerate(old)]
    N = [oracle.Cluster(list(m), i) for i, m in enumerate(new)]
    want = oracle.split_set_relation(oracle.SplitSet(O), oracle.SplitSet(N))
    assert got == want
    # member order after the call (in-place sorts of clusters > 20 members)
    for k in range(3):
        assert [so.clusters[k].members[i] for i in range(so.clusters[k].n_members)] == O[k].members
    for k in range(2):
        assert [sn.clusters[k].members[i] for i in range(sn.clusters[k].n_members)] == N[k].members


def test_relation_of_disjoint_split_sets(host, oracle):
    """Split sets that hold every point at most once (DBSCAN output): the
    connections and the member order left by the in-place sorts equal the
    oracle's pairwise restatement, sizes on both sides of 20 and ties of
    equal-sized old and new clusters included."""
    lib, _ = host
    rng = np.random.default_rng(7)
    for trial in range(12):
        n = int(rng.integers(200, 3000))
        perm = rng.permutation(n)
        sizes = rng.integers(1, 60, size=200)
        old, at = [], 0
        for k in sizes:
            if at + k > n * 0.9:
                break
            old.append(perm[at:at + k].tolist())
            at += int(k)
        # new clusters: unions of consecutive old clusters, part of a cluster
        # moved to the next one, some points dropped, one copy of an old
        # cluster (equal sizes)
        new, k = [], 0
        while k < len(old):
            step = int(rng.integers(1, 4))
            m = sum(old[k:k + step], [])
            if rng.random() < 0.3:
                m = m[: int(len(m) * rng.uniform(0.5, 1.0))]
            new.append(rng.permutation(m).tolist() if trial % 2 else m)
            k += step
        if trial % 3 == 0 and new:
            new[0] = list(old[0])
        new = [m for m in new if m]
        O = [oracle.Cluster(list(m), i) for i, m in enumerate(old)]
        N = [oracle.Cluster(list(m), i) for i, m in enumerate(new)]
        want = oracle.split_set_relation(oracle.SplitSet(O), oracle.SplitSet(N))
        so, ko = _mk_split([(m, i) for i, m in enumerate(old)])
        sn, kn = _mk_split([(m, i) for i, m in enumerate(new)])
        conn = lib.generate_split_set_relation(so, sn)
        got = [[conn[i].connections[k] for k in range(conn[i].n_connections)] for i in range(len(new))]
        assert got == want, trial
        for k in range(len(old)):
            assert [so.clusters[k].members[i] for i in range(so.clusters[k].n_members)] == O[k].members
        for k in range(len(new)):
            assert [sn.clusters[k].members[i] for i in range(sn.clusters[k].n_members)] == N[k].members


def test_dbscan_over_csr_matches_oracle(pkg, oracle):
    """mnhn_dbscan_csr (host code only) over epsilon lists built on the CPU
    from the oracle's distance matrix gives the oracle's clusters, member
    order included, for every epsilon up to the lists' bound."""
    path = os.path.join(GOLDEN, "inputs", "c2_small.fasta")
    seqs = oracle.read_fasta(path)[:160]
    dist = oracle.sw_matrix(seqs)
    n = len(seqs)
    eps_max = 500
    keep = (dist <= eps_max) & ~np.eye(n, dtype=bool)
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    col = np.nonzero(keep)[1].astype(np.int32)
    d = dist[keep].astype(np.uint16)
    for eps, minpts in ((100, 3), (250.5, 5), (400, 2), (500, 10), (0, 0)):
        got = pkg.dbscan_csr(rp, col, d, eps_max, eps, minpts)
        want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_matrix(dist))
        assert got == [c.members for c in want.clusters], (eps, minpts)


def test_ecsr_roundtrip_and_rejects_damaged_files(pkg, tmp_path):
    """.ecsr files (host code only): exact round trip, and truncated files,
    a wrong magic and out-of-range columns are refused."""
    rng = np.random.default_rng(5)
    n = 50
    counts = rng.integers(0, 7, n)
    rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(n, c, replace=False)) for c in counts]).astype(np.int32)
    dist = rng.integers(0, 3000, len(col)).astype(np.uint16)
    path = str(tmp_path / "a.ecsr")
    pkg.csr_save(path, rp, col, dist, 2999, 0x1234567890ABCDEF)
    r2, c2, d2, e2, h2 = pkg.csr_load(path)
    assert np.array_equal(r2, rp) and np.array_equal(c2, col) and np.array_equal(d2, dist)
    assert (e2, h2) == (2999, 0x1234567890ABCDEF)
    data = open(path, "rb").read()
    assert data[:8] == b"MNHNECSR" and len(data) == 40 + 8 * (n + 1) + 6 * len(col)
    bad = tmp_path / "b.ecsr"
    for blob in (data[:-1], b"X" + data[1:], data[:40 + 8 * (n + 1)] + struct.pack("<i", n) + data[44 + 8 * (n + 1):]):
        bad.write_bytes(blob)
        with pytest.raises(pkg.MnhnError):
            pkg.csr_load(str(bad))
    # empty lists
    pkg.csr_save(path, np.zeros(4, np.int64), np.zeros(0, np.int32), np.zeros(0, np.uint16), 10, 7)
    r3, c3, d3, e3, h3 = pkg.csr_load(path)
    assert len(r3) == 4 and len(c3) == 0 and (e3, h3) == (10, 7)


# ---- vector datasets (host code only) -----------------------------------------

class DataShape(ctypes.Structure):
    _fields_ = [("n_features", ctypes.c_size_t), ("n_samples", ctypes.c_size_t)]


def test_kmer_loader_matches_oracle(host, oracle):
    """shape_from_kmer_file / load_kmer_from_file_into_dataset read the
    reference's fasta2kmer files as the oracle does (dataset.c:346-425)."""
    lib, libc = host
    lib.shape_from_kmer_file.restype = DataShape
    lib.shape_from_kmer_file.argtypes = [ctypes.c_int]
    lib.load_kmer_from_file_into_dataset.restype = Dataset
    lib.load_kmer_from_file_into_dataset.argtypes = [ctypes.c_void_p, DataShape]
    lib.free_values_from_dataset.argtypes = [Dataset]
    libc.fdopen.restype = ctypes.c_void_p
    libc.fdopen.argtypes = [ctypes.c_int, ctypes.c_char_p]
    for fn in ("c1_small_k3.kmer", "c2_small_k4.kmer"):
        path = os.path.join(GOLDEN, "inputs", fn)
        fd = os.open(path, os.O_RDONLY)
        shape = lib.shape_from_kmer_file(fd)
        f = libc.fdopen(fd, b"r")
        ds = lib.load_kmer_from_file_into_dataset(f, shape)
        libc.fclose(f)
        want = oracle.read_kmers(path)
        assert (shape.n_features, shape.n_samples) == want.shape
        vals = ctypes.cast(ds.values, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)))
        got = np.array([[vals[k][i] for i in range(ds.n_values)] for k in range(ds.n_dimensions)],
                       np.float32)
        assert np.array_equal(got, want)
        lib.free_values_from_dataset(ds)


def test_values_cluster_files_match_oracle(host, oracle, tmp_path):
    """create_cluster_files_with_values (cluster_io.c:343-379): prefix-%03d
    files, "%f\\t" between values and "%f\\n" after the last."""
    lib, _ = host
    lib.create_cluster_files_with_values.argtypes = [ctypes.c_char_p, SplitSet, Dataset]
    rng = np.random.default_rng(8)
    dims, n = 3, 12
    v = (rng.normal(size=(dims, n)) * 1000).astype(np.float32)
    v[1, 3] = -0.0
    v[2, 5] = 1e-8
    rows = [np.ascontiguousarray(v[k]) for k in range(dims)]
    ptrs = (ctypes.POINTER(ctypes.c_float) * dims)(
        *[r.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for r in rows])
    ds = Dataset(None, None, None, n, ctypes.cast(ptrs, ctypes.c_void_p), dims, 0)
    cl = [([3, 1, 2], 0), ([5, 11, 0, 4], 1)]
    s, keep = _mk_split(cl)
    prefix = str(tmp_path / "vals")
    lib.create_cluster_files_with_values(prefix.encode(), s, ds)
    for i, (m, cid) in enumerate(cl):
        got = open("%s-%03d" % (prefix, i), "rb").read()
        assert got == oracle.values_file_text(oracle.Cluster(m, cid), v)


def test_dbscan_over_vcsr_matches_oracle(pkg, oracle):
    """mnhn_dbscan_vcsr (host code only) over vector lists built on the CPU
    from the oracle's accumulators at a larger threshold gives the oracle's
    clusters for every epsilon under it, L1 and L2."""
    values = oracle.read_kmers(os.path.join(GOLDEN, "inputs", "c1_small_k3.kmer"))
    n = values.shape[1]
    for norm, e_max, passes in ((1, 70.0, ((40, 5), (55.5, 3), (70, 8), (-2.0, 0))),
                                (2, 12.0, ((8, 5), (10.25, 4), (12, 2), (-9.0, 5)))):
        acc = oracle.vec_acc(values, norm)
        t_max = oracle.vec_threshold(norm, e_max)
        keep = acc <= t_max
        rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
        col = np.nonzero(keep)[1].astype(np.int32)
        val = acc[keep].astype(np.float32)
        for eps, minpts in passes:
            if oracle.vec_threshold(norm, eps) > t_max:
                continue
            got = pkg.dbscan_vcsr(rp, col, val, norm, t_max, eps, minpts)
            want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_acc(acc, norm))
            assert got == [c.members for c in want.clusters], (norm, eps, minpts)


class KmerFreq(ctypes.Structure):
    _fields_ = [("frequencies", ctypes.POINTER(ctypes.POINTER(ctypes.c_short))),
                ("n_kmers", ctypes.c_size_t), ("n_seq", ctypes.c_size_t)]


def test_write_kmer_base_matches_oracle(host, oracle, tmp_path):
    """write_kmer_base (kmers.c:378-389), host only: the hand formatter gives
    printf's "%hi" text, extremes and negative (wrapped) counts included."""
    lib, libc = host
    lib.write_kmer_base.argtypes = [ctypes.c_void_p, KmerFreq]
    rng = np.random.default_rng(2)
    for n_seq, nk in ((3, 4), (5, 64), (2, 1)):
        tab = rng.integers(-32768, 32768, size=(n_seq, nk)).astype(np.int16)
        tab[0, 0] = -32768
        tab[-1, -1] = 32767
        tab[0, -1] = 0
        rows = (ctypes.POINTER(ctypes.c_short) * n_seq)(
            *[tab[i].ctypes.data_as(ctypes.POINTER(ctypes.c_short)) for i in range(n_seq)])
        path = str(tmp_path / "k.txt")
        f = libc.fopen(path.encode(), b"w")
        lib.write_kmer_base(f, KmerFreq(rows, nk, n_seq))
        libc.fclose(f)
        assert open(path, "rb").read() == oracle.kmer_text(tab)


def test_kmer_column_counts(pkg):
    lib = pkg.load_library()
    lib.number_of_kmers_from_length.restype = ctypes.c_size_t
    lib.number_of_kmers_from_length.argtypes = [ctypes.c_size_t]
    lib.number_of_protein_kmers_from_length.restype = ctypes.c_size_t
    lib.number_of_protein_kmers_from_length.argtypes = [ctypes.c_size_t]
    for k in range(1, 12):
        assert lib.number_of_kmers_from_length(k) == 4 ** k
        assert lib.number_of_protein_kmers_from_length(k) == 20 ** k
    assert [lib.mnhn_kmer_columns(k, 0) for k in (0, 1, 15, 16)] == [0, 4, 4 ** 15, 0]
    assert [lib.mnhn_kmer_columns(k, 1) for k in (0, 1, 7, 8)] == [0, 20, 20 ** 7, 0]


def test_dbscan_threaded_pass_lists_match_oracle(pkg, oracle, monkeypatch):
    """Lists large enough (> 2^20 stored entries) that each pass's neighbour
    lists are built by several host threads: clusters equal the oracle's for
    Smith-Waterman-style lists (self inserted) and vector lists (self stored)."""
    monkeypatch.setenv("MNHN_HOST_THREADS", "7")
    rng = np.random.default_rng(12)
    n = 3000
    centre = rng.integers(0, 60, n)
    d = np.abs(centre[:, None] - centre[None, :]) * 40 + rng.integers(0, 120, (n, n))
    d = np.triu(d, 1)
    d = (d + d.T).astype(np.int64)
    eps_max = 2000
    keep = (d <= eps_max) & ~np.eye(n, dtype=bool)
    assert keep.sum() > (1 << 20)
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    col = np.nonzero(keep)[1].astype(np.int32)
    dist = d[keep].astype(np.uint16)
    for eps, minpts in ((60, 4), (150.5, 10), (-2.0, 3)):
        got = pkg.dbscan_csr(rp, col, dist, eps_max, eps, minpts)
        want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_matrix(d))
        assert got == [c.members for c in want.clusters], (eps, minpts)
    acc = d.astype(np.float32)
    keepv = acc <= np.float32(eps_max)
    rpv = np.concatenate([[0], np.cumsum(keepv.sum(1))]).astype(np.int64)
    colv = np.nonzero(keepv)[1].astype(np.int32)
    val = acc[keepv]
    for eps, minpts in ((60.0, 4), (140.0, 12)):
        got = pkg.dbscan_vcsr(rpv, colv, val, 1, float(eps_max), eps, minpts)
        want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_acc(acc, 1))
        assert got == [c.members for c in want.clusters], (eps, minpts)


def test_dbscan_dense_neighbourhoods_match_oracle(pkg, oracle):
    """Neighbourhoods covering most points: the expansion walks the points
    still outside the cluster and looks them up in N(q) instead of scanning
    N(q); the clusters and member order equal the oracle's."""
    rng = np.random.default_rng(21)
    n = 1200
    centre = rng.integers(0, 8, n)
    d = np.abs(centre[:, None] - centre[None, :]) * 100 + rng.integers(0, 300, (n, n))
    d = np.triu(d, 1)
    d = (d + d.T).astype(np.int64)
    keep = (d <= 5000) & ~np.eye(n, dtype=bool)
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    col = np.nonzero(keep)[1].astype(np.int32)
    dist = d[keep].astype(np.uint16)
    for eps, minpts in ((900, 5), (600, 50), (400, 5), (1500, 2)):
        got = pkg.dbscan_csr(rp, col, dist, 5000, eps, minpts)
        want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_matrix(d))
        assert got == [c.members for c in want.clusters], (eps, minpts)


def _lists_from_matrix(d, bound):
    keep = (d <= bound) & ~np.eye(len(d), dtype=bool)
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    return rp, np.nonzero(keep)[1].astype(np.int32), d[keep].astype(np.uint16)


def test_dbscan_sweep_matches_oracle(pkg, oracle, monkeypatch):
    """mnhn_dbscan_csr_sweep: passes over lists ordered by distance once give
    the oracle's clusters, member order included, for every epsilon; rows
    cover the three orderings (short rows, integer ranges short enough for
    one counting pass, wide ranges by radix), large clusters that find the
    points outside them through those points' short lists (symmetric lists)
    or by binary search (not declared symmetric), and lists big enough for
    the threaded build."""
    seqs = oracle.read_fasta(os.path.join(GOLDEN, "inputs", "c2_small.fasta"))[:160]
    dist = oracle.sw_matrix(seqs)
    rp, col, d = _lists_from_matrix(dist, 500)
    eps = [100, 250.5, 400, 500, 0, -2.0, 330]
    for minpts in (3, 0):
        got = pkg.dbscan_csr_sweep(rp, col, d, 500, eps, minpts)
        for e, g in zip(eps, got):
            want = oracle.dbscan(len(seqs), e, minpts, oracle.region_query_from_matrix(dist))
            assert g == [c.members for c in want.clusters], (e, minpts)

    rng = np.random.default_rng(31)
    n = 1500
    centre = rng.integers(0, 8, n)
    m = np.abs(centre[:, None] - centre[None, :]) * 100 + rng.integers(0, 300, (n, n))
    wide = rng.random(n) < 0.3  # rows whose far entries spread over 0..30000
    m = np.where(wide[:, None] | wide[None, :], m * 6 + rng.integers(0, 9000, (n, n)), m)
    m = np.triu(m, 1)
    m = (m + m.T).astype(np.int64)
    rp, col, d = _lists_from_matrix(m, 30000)
    eps = [900, 600, 400, 1500, 3000, 20000]
    for minpts in (5, 50, 2):
        got = pkg.dbscan_csr_sweep(rp, col, d, 30000, eps, minpts)
        for e, g in zip(eps, got):
            want = oracle.dbscan(n, e, minpts, oracle.region_query_from_matrix(m))
            assert g == [c.members for c in want.clusters], (e, minpts)
        # without the symmetric declaration (no triggers): the same clusters
        assert pkg.dbscan_csr_sweep(rp, col, d, 30000, eps, minpts, symmetric=False) == got

    monkeypatch.setenv("MNHN_HOST_THREADS", "7")
    n = 3000
    centre = rng.integers(0, 60, n)
    m = np.abs(centre[:, None] - centre[None, :]) * 40 + rng.integers(0, 120, (n, n))
    m = np.triu(m, 1)
    m = (m + m.T).astype(np.int64)
    rp, col, d = _lists_from_matrix(m, 2000)
    assert len(col) > (1 << 20)
    eps = [60, 150.5, 1000, 2000]
    got = pkg.dbscan_csr_sweep(rp, col, d, 2000, eps, 4)
    for e, g in zip(eps, got):
        want = oracle.dbscan(n, e, 4, oracle.region_query_from_matrix(m))
        assert g == [c.members for c in want.clusters], e
    # the same passes without the index (MNHN_SWEEP_INDEX=0) agree
    monkeypatch.setenv("MNHN_SWEEP_INDEX", "0")
    assert pkg.dbscan_csr_sweep(rp, col, d, 2000, eps, 4) == got


def test_dbscan_vcsr_sweep_matches_oracle(pkg, oracle):
    """mnhn_dbscan_vcsr_sweep over vector lists (self stored, float keys,
    infinities and exact zeros among them) gives the oracle's clusters for
    every epsilon, L1 and L2."""
    values = oracle.read_kmers(os.path.join(GOLDEN, "inputs", "c1_small_k3.kmer"))
    n = values.shape[1]
    for norm, e_max, eps in ((1, 70.0, [40, 55.5, 70, -2.0, 0.0]), (2, 12.0, [8, 10.25, 12, -9.0, 0.0])):
        acc = oracle.vec_acc(values, norm)
        t_max = oracle.vec_threshold(norm, e_max)
        keep = acc <= t_max
        rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
        col = np.nonzero(keep)[1].astype(np.int32)
        for minpts in (5, 0):
            got = pkg.dbscan_vcsr_sweep(rp, col, acc[keep].astype(np.float32), norm, t_max, eps, minpts)
            for e, g in zip(eps, got):
                want = oracle.dbscan(n, e, minpts, oracle.region_query_from_acc(acc, norm))
                assert g == [c.members for c in want.clusters], (norm, e, minpts)
    rng = np.random.default_rng(3)
    pts = rng.integers(0, 4, size=(300, 6)).astype(np.float32)
    pts[7, 2] = np.inf
    pts[50:60] = pts[49]
    acc = oracle.vec_acc(pts.T, 1)
    keep = acc <= np.float32(np.inf)
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    col = np.nonzero(keep)[1].astype(np.int32)
    eps = [0.0, 2.0, 5.0, float(np.inf)]
    got = pkg.dbscan_vcsr_sweep(rp, col, acc[keep].astype(np.float32), 1, float(np.inf), eps, 3)
    for e, g in zip(eps, got):
        want = oracle.dbscan(300, e, 3, oracle.region_query_from_acc(acc, 1))
        assert g == [c.members for c in want.clusters], e


def test_dbscan_lists_with_distances_beyond_32767(pkg, oracle, tmp_path):
    """List distances are unsigned 16-bit: pairs with d in [32768, 65535]
    (sequences of 8,192 bp and more) keep their epsilon test and their place
    in the distance order, in single passes, sweeps and .ecsr files.  A signed
    16-bit distance would turn every such pair into a neighbour at any
    epsilon and put it first in the distance order."""
    rng = np.random.default_rng(41)
    n = 500
    centre = rng.integers(0, 12, n)
    d = np.abs(centre[:, None] - centre[None, :]) * 2700 + 30000 + rng.integers(0, 1500, (n, n))
    d = np.triu(d, 1)
    d = (d + d.T).astype(np.int64)
    assert d.max() <= 65535 and (d >= 32768).mean() > 0.5
    rp, col, dist = _lists_from_matrix(d, 65535)
    assert dist.dtype == np.uint16 and int(dist.max()) == int(d.max())
    eps = [31000.0, 31500.5, 34000, 40000, 65535]
    for minpts in (3, 12):
        for e in eps:
            want = oracle.dbscan(n, e, minpts, oracle.region_query_from_matrix(d))
            assert pkg.dbscan_csr(rp, col, dist, 65535, e, minpts) == [c.members for c in want.clusters], (e, minpts)
        got = pkg.dbscan_csr_sweep(rp, col, dist, 65535, eps, minpts)
        for e, g in zip(eps, got):
            want = oracle.dbscan(n, e, minpts, oracle.region_query_from_matrix(d))
            assert g == [c.members for c in want.clusters], (e, minpts)
    path = str(tmp_path / "wide.ecsr")
    pkg.csr_save(path, rp, col, dist, 65535, 99)
    r2, c2, d2, e2, _ = pkg.csr_load(path)
    assert np.array_equal(r2, rp) and np.array_equal(c2, col) and np.array_equal(d2, dist) and e2 == 65535


def test_dbscan_sweep_long_lists_outside_a_cluster(pkg, oracle):
    """Points on a line with a sparse tail: when the cluster has almost
    everything, the points still outside have long lists (binary search) or
    short ones (found from their own lists), both merged in ascending order;
    clusters equal the oracle's."""
    pos = np.concatenate([np.arange(4000), 3999 + 50 * np.arange(1, 31)]).astype(np.int64)
    n = len(pos)
    eps_max = 1800
    rows = [np.nonzero(np.abs(pos - pos[p]) <= eps_max)[0] for p in range(n)]
    rows = [r[r != p] for p, r in enumerate(rows)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    col = np.concatenate(rows).astype(np.int32)
    d = np.abs(pos[np.repeat(np.arange(n), np.diff(rp))] - pos[col]).astype(np.uint16)

    def rq(p, e):
        lo = np.searchsorted(pos, pos[p] - np.float32(e), side="left")
        hi = np.searchsorted(pos, pos[p] + np.float32(e), side="right")
        return list(range(lo, hi))

    eps = [1800, 1200, 700]
    got = pkg.dbscan_csr_sweep(rp, col, d, eps_max, eps, 4)
    for e, g in zip(eps, got):
        want = oracle.dbscan(n, e, 4, rq)
        assert g == [c.members for c in want.clusters], e
    assert pkg.dbscan_csr_sweep(rp, col, d, eps_max, eps, 4, symmetric=False) == got


def test_dbscan_passes_over_whole_lists_match_oracle(pkg, oracle):
    """A pass whose epsilon keeps every stored entry reads the rows in place
    (no filtered copy): Smith-Waterman lists (self inserted) and vector lists
    (self stored), dense neighbourhoods included, give the oracle's clusters."""
    rng = np.random.default_rng(23)
    n = 1400
    centre = rng.integers(0, 8, n)
    d = np.abs(centre[:, None] - centre[None, :]) * 100 + rng.integers(0, 300, (n, n))
    d = np.triu(d, 1)
    d = (d + d.T).astype(np.int64)
    for eps, minpts in ((900, 5), (400, 5), (1500, 2), (250, 40)):
        rp, col, dist = _lists_from_matrix(d, eps)
        got = pkg.dbscan_csr(rp, col, dist, eps, eps, minpts)
        want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_matrix(d))
        assert got == [c.members for c in want.clusters], (eps, minpts)
        acc = d.astype(np.float32)
        keep = acc <= np.float32(eps)
        rpv = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
        colv = np.nonzero(keep)[1].astype(np.int32)
        got = pkg.dbscan_vcsr(rpv, colv, acc[keep], 1, float(eps), eps, minpts)
        want = oracle.dbscan(n, eps, minpts, oracle.region_query_from_acc(acc, 1))
        assert got == [c.members for c in want.clusters], ("L1", eps, minpts)


def test_dbscan_parallel_row_scans_match_oracle(pkg, oracle, monkeypatch):
    """Lists long enough (> 4M entries) that host threads scan batches of
    queue rows ahead of the expansion: materialised pass lists, rows read in
    place, and the sweep's distance-ordered rows all give the oracle's
    clusters, member order included."""
    monkeypatch.setenv("MNHN_HOST_THREADS", "7")
    pos = np.arange(6000, dtype=np.int64)
    pos[2500:] += 900   # a gap wider than every epsilon below: two clusters at least
    pos[4200:] += 350   # a gap only the larger epsilons bridge
    n = len(pos)
    eps_max = 700
    rows = [np.nonzero(np.abs(pos - pos[p]) <= eps_max)[0] for p in range(n)]
    rows = [r[r != p] for p, r in enumerate(rows)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    col = np.concatenate(rows).astype(np.int32)
    d = np.abs(pos[np.repeat(np.arange(n), np.diff(rp))] - pos[col]).astype(np.uint16)
    assert len(col) > (1 << 22)

    def rq(p, e):
        lo = np.searchsorted(pos, pos[p] - np.float32(e), side="left")
        hi = np.searchsorted(pos, pos[p] + np.float32(e), side="right")
        return list(range(lo, hi))

    eps = [700, 500, 300]
    want = {e: [c.members for c in oracle.dbscan(n, e, 6, rq).clusters] for e in eps}
    for e in eps:
        assert pkg.dbscan_csr(rp, col, d, eps_max, e, 6) == want[e], e
    for sym in (True, False):
        got = pkg.dbscan_csr_sweep(rp, col, d, eps_max, eps, 6, symmetric=sym)
        assert got == [want[e] for e in eps], sym
