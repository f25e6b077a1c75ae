"""CPU restatement of the reference's DBSCAN / adaptive-clustering path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, and only as the checker.  The shipped library
never imports it.

Every function names the reference lines it restates (paths are relative to
the reference repository root).  Distances come either from the C restatement
of smith-waterman.c (oracle/liboracle_sw.so, via ``sw_matrix``) or from a
caller-supplied matrix, so the clustering logic here can be checked against
golden files written by the compiled reference and then used to check the
product's clustering on inputs too large for the reference to run.
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    """Load oracle/liboracle_sw.so (built by ``make -C oracle oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_sw.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle oracle`")
        lib = ctypes.CDLL(path)
        lib.oracle_sw_distance.restype = ctypes.c_uint
        lib.oracle_sw_distance.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.c_size_t, ctypes.c_size_t]
        lib.oracle_sw_pairs.restype = ctypes.c_int
        lib.oracle_sw_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_int]
        lib.oracle_sw_matrix.restype = ctypes.c_int
        lib.oracle_sw_matrix.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_int]
        lib.oracle_vec_rows.restype = ctypes.c_int
        lib.oracle_vec_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int]
        lib.oracle_cpu_allpairs.restype = ctypes.c_double
        lib.oracle_cpu_allpairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_uint64)]
        _LIB = lib
    return _LIB


class _SeqArgs:
    """Keeps C views (char**, size_t*) of a list of byte strings alive."""

    def __init__(self, seqs: Sequence[bytes]):
        self.bufs = [ctypes.create_string_buffer(s, len(s) + 1) for s in seqs]
        self.ptrs = (ctypes.c_char_p * len(seqs))(
            *[ctypes.cast(b, ctypes.c_char_p) for b in self.bufs])
        self.lens = (ctypes.c_size_t * len(seqs))(*[len(s) for s in seqs])


def sw_distance(a: bytes, b: bytes) -> int:
    """smith-waterman.c:16-81 -- one pair."""
    return int(_lib().oracle_sw_distance(a, b, len(a), len(b)))


def sw_pairs(seqs: Sequence[bytes], ij: np.ndarray, nthreads: int = 8) -> np.ndarray:
    """d(seqs[i], seqs[j]) for each row (i, j) of ij."""
    ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1, 2)
    out = np.zeros(len(ij), dtype=np.int32)
    sa = _SeqArgs(seqs)
    _lib().oracle_sw_pairs(sa.ptrs, sa.lens, ij.ctypes.data, len(ij),
                           out.ctypes.data, nthreads)
    return out


def sw_matrix(seqs: Sequence[bytes], nthreads: int = 8) -> np.ndarray:
    """Full N x N distance matrix (unique pairs computed once, mirrored)."""
    n = len(seqs)
    out = np.zeros((n, n), dtype=np.int32)
    sa = _SeqArgs(seqs)
    _lib().oracle_sw_matrix(sa.ptrs, sa.lens, n, out.ctypes.data, nthreads)
    return out


def cpu_allpairs(seqs: Sequence[bytes], nthreads: int, ordered: bool):
    """Time the restated SW over all pairs: returns (seconds, checksum)."""
    sa = _SeqArgs(seqs)
    cs = ctypes.c_uint64(0)
    t = _lib().oracle_cpu_allpairs(sa.ptrs, sa.lens, len(seqs), nthreads,
                                   1 if ordered else 0, None, ctypes.byref(cs))
    return float(t), int(cs.value)


# --------------------------------------------------------------------------
# FASTA reader: dataset.c:264-344
# --------------------------------------------------------------------------

def read_fasta(path: str) -> List[bytes]:
    """dataset_from_fasta: a '>' line opens a record; any other line adds its
    first whitespace-delimited token (<2000 chars), uppercased.  A line with
    no token re-adds the previous token, because sscanf leaves the line
    buffer unchanged (dataset.c:316)."""
    seqs: List[bytearray] = []
    last = b""
    with open(path, "rb") as f:
        for line in f:
            if line[:1] == b">":
                seqs.append(bytearray())
                continue
            toks = line.split()
            if toks:
                last = toks[0].upper()
            if not seqs:
                continue  # text before the first header is undefined there
            seqs[-1] += last
    return [bytes(s) for s in seqs]


# --------------------------------------------------------------------------
# DBSCAN: dbscan.c:783-1047
# --------------------------------------------------------------------------

@dataclass
class Cluster:
    members: List[int]
    id: int


@dataclass
class SplitSet:
    clusters: List[Cluster] = field(default_factory=list)

    @property
    def n_clusters(self) -> int:
        return len(self.clusters)


def region_query_from_matrix(dist: np.ndarray) -> Callable[[int, float], List[int]]:
    """region_query (dbscan.c:525-561): ascending j with (float)d <= eps."""
    df = dist.astype(np.float32)

    def rq(point: int, eps: float) -> List[int]:
        return np.nonzero(df[point] <= np.float32(eps))[0].tolist()

    return rq


def dbscan(n: int, eps: float, minpts: int,
           region_query: Callable[[int, float], List[int]],
           out_lines: Optional[List[str]] = None) -> SplitSet:
    """dbscan_SW (dbscan.c:950-1047) with expand_cluster (dbscan.c:783-948).

    The queue append of expand_cluster (dbscan.c:892-912) is restated with its
    intended meaning: the members of N(q) not yet in the sorted union, in
    ascending order."""
    eps = float(np.float32(eps))
    if eps == -1:  # dbscan.c:980-984
        if out_lines is not None:
            out_lines.append("Warning @dbscan: epsilon value is: %f\n" % eps)
        eps = float(np.float32(0.015))
        if out_lines is not None:
            out_lines.append("Warning @dbscan: epsilon was changed to: %f\n" % eps)
    if minpts == -1:  # dbscan.c:985-989
        if out_lines is not None:
            out_lines.append("Warning @dbscan: minpts value is: %i\n" % minpts)
        minpts = 5
        if out_lines is not None:
            out_lines.append("Warning @dbscan: minpts was changed to: %i\n" % minpts)
    visited = np.zeros(n, dtype=bool)
    member = np.zeros(n, dtype=bool)
    clusters: List[Cluster] = []
    for i in range(n):
        if visited[i]:
            continue
        visited[i] = True
        nb = region_query(i, eps)
        if minpts < len(nb):  # strict seed test, dbscan.c:1011
            cl = Cluster(members=[i], id=len(clusters))
            member[i] = True
            union = sorted(nb)
            in_union = set(union)
            queue = list(nb)
            k = 0
            while k < len(queue):
                q = queue[k]
                if not visited[q]:
                    visited[q] = True
                    nbq = region_query(q, eps)
                    if len(nbq) >= minpts:  # expansion test, dbscan.c:847
                        for x in nbq:
                            if x not in in_union:
                                in_union.add(x)
                                queue.append(x)
                if not member[q]:  # dbscan.c:934-941
                    cl.members.append(q)
                    member[q] = True
                k += 1
            clusters.append(cl)
    return SplitSet(clusters)


# --------------------------------------------------------------------------
# split-set relations / coverage / formats: cluster_io.c
# --------------------------------------------------------------------------

def data_not_in_clusters(s: SplitSet, n: int) -> int:
    """cluster_io.c:564-593 -- count of points in no cluster."""
    covered = np.zeros(n, dtype=bool)
    for c in s.clusters:
        covered[c.members] = True
    return int(n - covered.sum())


def intersection_count(a: Cluster, b: Cluster) -> int:
    """intersection_of_clusters (cluster_io.c:447-501), including its side
    effect: the larger cluster (a on ties) is sorted in place when it has
    more than 20 members (cluster_io.c:467-469)."""
    if len(a.members) >= len(b.members):
        big, small = a, b
    else:
        big, small = b, a
    if len(big.members) > 20:
        big.members.sort()
    bs = set(big.members)
    return sum(1 for x in small.members if x in bs)


def split_set_relation(ancient: SplitSet, new: SplitSet) -> List[List[int]]:
    """generate_split_set_relation (cluster_io.c:407-445): new cluster i is
    connected to ancient cluster j when |A_j & N_i| > 0.8 |A_j|."""
    conns = []
    for nc in new.clusters:
        row = []
        for j, ac in enumerate(ancient.clusters):
            inter = intersection_count(ac, nc)
            if float(inter) > float(len(ac.members)) * 0.8:
                row.append(j)
        conns.append(row)
    return conns


def split_set_bytes(s: SplitSet) -> bytes:
    """store_split_set (cluster_io.c:638-669) byte layout, little endian."""
    out = bytearray(b"SPLIT_SET")
    out += struct.pack("<i", s.n_clusters)
    for c in s.clusters:
        out += struct.pack("<ii", len(c.members), c.id)
        out += np.asarray(c.members, dtype="<i4").tobytes()
    return bytes(out)


def cluster_file_text(cl: Cluster, seqs: Sequence[bytes]) -> bytes:
    """create_single_cluster_file (cluster_io.c:672-699): '>sequence_<idx>'
    header and a newline before every 50th residue after the first."""
    out = bytearray()
    for m in cl.members:
        out += b">sequence_%d\n" % m
        s = seqs[m]
        for k in range(len(s)):
            if k != 0 and k % 50 == 0:
                out += b"\n"
            out += s[k:k + 1]
        out += b"\n"
    return bytes(out)


def cluster_dbscan_sw(seqs: Sequence[bytes], eps: float, minpts: int,
                      dist: Optional[np.ndarray] = None):
    """cluster_dbscan_SW.c:25-71: returns (stdout text, {suffix: bytes})."""
    if dist is None:
        dist = sw_matrix(seqs)
    lines: List[str] = []
    s = dbscan(len(seqs), eps, minpts, region_query_from_matrix(dist), lines)
    lines.append("%u clusters obtained \n" % s.n_clusters)
    files = {}
    if s.n_clusters < 500:
        for i, c in enumerate(s.clusters):
            files["-%03d" % i] = cluster_file_text(c, seqs)
    return "".join(lines), files, s


def adaptive_clustering(seqs: Sequence[bytes], eps_start: float, eps_inc: float,
                        minpts: int, dist: Optional[np.ndarray] = None,
                        max_passes: int = 100000,
                        rq: Optional[Callable[[int, float], List[int]]] = None,
                        n: Optional[int] = None):
    """adaptive_dbscan (dbscan.c:1117-1380) with one worker thread, and the
    main() prints of adaptive_clustering.c.  Returns (stdout text,
    {suffix: split-set bytes}, kept layers) or raises on the reference's
    _exit(1) paths.  rq / n replace the Smith-Waterman region query (e.g.
    region_query_from_acc for the L1 / L2 builds)."""
    if rq is None:
        if dist is None:
            dist = sw_matrix(seqs)
        rq = region_query_from_matrix(dist)
    if n is None:
        n = len(seqs)
    f32 = np.float32
    eps_start = f32(eps_start)
    eps_inc = f32(eps_inc)
    out: List[str] = ["Dataset read!\n"]
    out.append("Performing adaptive clustering with parameters: \n"
               "minpts: %i\nepsilon_start: %f\nepsilon_inc: %f\n"
               % (minpts, float(eps_start), float(eps_inc)))
    s0 = dbscan(n, float(eps_start), minpts, rq, out)
    out.append("Initial set obtained with %i clusters\n" % s0.n_clusters)
    tries = 0
    while s0.n_clusters == 1:
        if tries == 20 or eps_start == 0:
            out.append("Error did not find a sufficient starting position in 20 tries \n")
            raise RuntimeError("".join(out))
        eps_start = f32(eps_start / f32(2))
        out.append("Trying new starting point \n")
        s0 = dbscan(n, float(eps_start), minpts, rq, out)
        tries += 1
    out.append("Sarting with %i clusters\n" % s0.n_clusters)
    nc = data_not_in_clusters(s0, n)
    out.append("Coverage at initial point: %f\n"
               % float(f32(1) - f32(nc) / f32(n)))
    # worker loop (dbscan.c:1061-1115): eps_now starts at eps_start + inc
    sets = [s0]
    eps_now = f32(eps_start + eps_inc)
    passes = 0
    while True:
        eps = eps_now
        eps_now = f32(eps_now + eps_inc)
        s = dbscan(n, float(eps), minpts, rq, out)
        sets.append(s)
        passes += 1
        if s.n_clusters == 1:
            break
        if passes >= max_passes:
            raise RuntimeError("no single-cluster pass within max_passes")
    stop = len(sets) - 1
    kept = [s0]
    conns = []
    for i in range(1, stop + 1):
        ns = sets[i]
        if ns.n_clusters < kept[-1].n_clusters:
            count = len(kept)
            nc = data_not_in_clusters(ns, n)
            out.append("Layer %i has %i clusters\n" % (count, ns.n_clusters))
            out.append("  Coverage at layer %i: %f\n"
                       % (count, float(f32(1) - f32(nc) / f32(n))))
            out.append("  Epsilon at layer %i: %f\n"
                       % (count, float(f32(eps_start + f32(i + 1) * eps_inc))))
            conns.append(split_set_relation(kept[-1], ns))
            kept.append(ns)
    out.append("Connections found: \n")
    files = {"%04d" % i: split_set_bytes(s) for i, s in enumerate(kept)}
    for i in range(len(kept) - 1, 0, -1):
        out.append("Layer %i:\n" % i)
        for j in range(kept[i].n_clusters):
            out.append("  Cluster %i connected to: \n  " % j)
            out.append("".join("%i " % c for c in conns[i - 1][j]))
            out.append("\n")
    return "".join(out), files, kept


# --------------------------------------------------------------------------
# the other score() consumers: find_closest_sequence_SW.c, filter.c,
# comparison.c, dataset.c:85-110
# --------------------------------------------------------------------------

def _cross(a: Sequence[bytes], b: Sequence[bytes]) -> np.ndarray:
    """d(a_i, b_j) as an int64 matrix (oracle distances)."""
    seqs = list(a) + list(b)
    ij = np.array([[i, len(a) + j] for i in range(len(a)) for j in range(len(b))], np.int32)
    if len(ij) == 0:
        return np.zeros((len(a), len(b)), np.int64)
    return sw_pairs(seqs, ij).astype(np.int64).reshape(len(a), len(b))


def find_closest(db: Sequence[bytes], query: bytes):
    """find_closest_sequence_SW.c:26-63, 126-153 (query index 1): the smallest
    distance and the lowest index that has it; stdout text and output file."""
    d = _cross([query], db)[0]
    best = int(np.argmin(d))
    text = "Smith Waterman distance = %i \n" % int(d[best])
    return text, cluster_file_text(Cluster([best], 0), db)


def dataset_to_fasta(seqs: Sequence[bytes]) -> bytes:
    """dataset.c:85-110 for non-empty sequences."""
    return cluster_file_text(Cluster(list(range(len(seqs))), 0), seqs) if seqs else b""


def find_satellite(ds: Sequence[bytes], target: bytes, size: int, plus: int, minus: int,
                   max_distance: int) -> bytes:
    """find_satellite.c:12-74 with filter.c:18-140: the size window in size_t
    arithmetic, then d(seq, target) <= max_distance, printed in order."""
    m64 = (1 << 64) - 1
    hi, lo = (size + plus) & m64, (size - minus) & m64
    sized = [s for s in ds if lo <= len(s) <= hi]
    if not sized:
        return b""
    d = _cross(sized, [target])[:, 0]
    return dataset_to_fasta([s for s, x in zip(sized, d) if max_distance >= int(x)])


def _to_int_x86(t: float) -> int:
    """cvttsd2si: truncation toward zero, INT_MIN outside the int range."""
    if t != t or not (-2147483649.0 < t < 2147483648.0):
        return -2147483648
    return int(t)


def compare_sw(a: Sequence[bytes], b: Sequence[bytes]) -> str:
    """compareSW.c with comparison.c:113-277 run by one thread: the mean from
    an int sum of the distances (modulo 2^32), sigma from an int sum that
    takes each (d - mean)^2 through a double; rows then columns in order."""
    d = _cross(a, b).ravel()
    s32 = int(d.sum()) & 0xFFFFFFFF
    s32 = s32 - (1 << 32) if s32 >= (1 << 31) else s32
    mean = float(s32) / (float(len(a)) * float(len(b)))
    acc = 0
    for v in d.tolist():
        x = (float(v) - mean) * (float(v) - mean)
        acc = _to_int_x86(float(acc) + x)
    sigma = float(np.sqrt(np.float64(acc))) / (len(a) + len(b))
    return "mean = %s, sigma = %s \n" % (_fmt_lf(mean), _fmt_lf(sigma))


def silhouette(a: Sequence[bytes], b: Sequence[bytes]) -> str:
    """silhouette.c with comparison.c:20-75, 279-313 run by one thread."""
    internal = _cross(a, a).sum(1)
    external = _cross(a, b).sum(1)
    s_mean = 0.0
    for i in range(len(a)):
        av = float(internal[i]) / (float(len(a)) - 1.0)
        bv = float(external[i]) / float(len(b))
        if av < bv:
            s = 1 - av / bv
        elif av == bv:
            s = 0.0
        else:
            s = bv / av - 1
        s_mean += s
    return "%s\n" % _fmt_lf(s_mean / float(len(a)))


def _fmt_lf(x: float) -> str:
    """printf("%lf") as glibc prints it (nan keeps its sign)."""
    if x != x:
        return "-nan" if np.signbit(x) else "nan"
    return "%f" % x


def compare_norms(seqs: Sequence[bytes], proj: np.ndarray) -> str:
    """compare_norms.c + print_SW_PCA_L_comparison (comparison.c:343-387):
    header, then per pair j < i the distance, the L2 and the L1 norm of the
    float32 projection differences (each difference taken in float32)."""
    n = len(seqs)
    d = sw_matrix(seqs) if n > 1 else np.zeros((n, n), np.int64)
    p32 = proj.astype(np.float32)
    out = ["i j SW L1 L2\n"]
    for i in range(n):
        for j in range(i):
            buf = (p32[j] - p32[i]).astype(np.float64)
            l2 = 0.0
            l1 = 0.0
            for x in buf.tolist():
                l2 += x * x
                l1 += abs(x)
            out.append("%i %i %s %s %s\n" % (i, j, _fmt_lf(float(d[i, j])), _fmt_lf(float(np.sqrt(l2))),
                                             _fmt_lf(l1)))
    return "".join(out)


# --------------------------------------------------------------------------
# L1 / L2 vector scans: dbscan.c:563-780 (-D_SCAN_L1 / -D_SCAN_L2),
# dataset.c:346-469 (k-mer and projection files), cluster_io.c:343-379
# --------------------------------------------------------------------------

_STRTOF = None


def _strtof(tok: bytes) -> np.float32:
    """libc strtof, the conversion fscanf("%f") performs."""
    global _STRTOF
    if _STRTOF is None:
        f = ctypes.CDLL(None).strtof
        f.restype = ctypes.c_float
        f.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
        _STRTOF = f
    return np.float32(_STRTOF(tok, None))


def read_projections(path: str, dims: int, n: int) -> np.ndarray:
    """load_projections_from_file_into_dataset (dataset.c:427-469): n rows of
    dims whitespace-separated floats; returns values[dim][row] (float32)."""
    toks = open(path, "rb").read().split()
    if len(toks) < n * dims:
        raise ValueError("Error reading projections file")
    v = np.array([_strtof(t) for t in toks[:n * dims]], np.float32)
    return np.ascontiguousarray(v.reshape(n, dims).T)


def kmer_shape(data: bytes):
    """shape_from_kmer_file (dataset.c:346-375): (features = tabs before the
    first line break, samples = line breaks from that one on)."""
    first = data.index(b"\n")
    return data[:first].count(b"\t"), data[first:].count(b"\n")


def read_kmers(path: str) -> np.ndarray:
    """load_kmer_from_file_into_dataset (dataset.c:378-425): per sample a
    label token and n_features floats; returns values[feature][sample]."""
    data = open(path, "rb").read()
    nf, ns = kmer_shape(data)
    toks = data.split()
    v = np.zeros((nf, ns), np.float32)
    k = 0
    for i in range(ns):
        k += 1  # label
        for j in range(nf):
            v[j, i] = _strtof(toks[k])
            k += 1
    return v


def vec_acc(values: np.ndarray, norm: int, rows: Optional[Sequence[int]] = None,
            nthreads: int = 8) -> np.ndarray:
    """Accumulators of the query rows against every point, in the order of
    region_query's AVX2 + FMA branch (vec_oracle.c); values is [dims][n]."""
    v = np.ascontiguousarray(values, dtype=np.float32)
    dims, n = v.shape
    rows = np.arange(n, dtype=np.int32) if rows is None else np.asarray(rows, np.int32)
    ptrs = (ctypes.c_void_p * max(dims, 1))(*[v[k].ctypes.data for k in range(dims)])
    out = np.empty((len(rows), n), np.float32)
    rc = _lib().oracle_vec_rows(ptrs, dims, n, norm, rows.ctypes.data, len(rows), out.ctypes.data, nthreads)
    if rc:
        raise ValueError("norm must be 1 or 2")
    return out


def vec_threshold(norm: int, eps: float) -> np.float32:
    """dbscan.c:577, 676-679: eps * eps in float32 for L2, eps for L1."""
    e = np.float32(eps)
    return np.float32(e * e) if norm == 2 else e


def region_query_from_acc(acc: np.ndarray, norm: int) -> Callable[[int, float], List[int]]:
    """region_query (dbscan.c:563-780): ascending x with acc(p, x) <= thr."""
    def rq(point: int, eps: float) -> List[int]:
        return np.nonzero(acc[point] <= vec_threshold(norm, eps))[0].tolist()

    return rq


def values_file_text(cl: Cluster, values: np.ndarray) -> bytes:
    """create_single_cluster_file_with_values (cluster_io.c:343-358)."""
    dims = values.shape[0]
    out = []
    for m in cl.members:
        row = [float(values[k, m]) for k in range(dims)]
        out.append("".join("%f\t" % x for x in row[:-1]) + "%f\n" % row[-1])
    return "".join(out).encode()


def cluster_dbscan_vec(seqs: Sequence[bytes], values: np.ndarray, norm: int, eps: float,
                       minpts: int, with_values: bool = False):
    """cluster_dbscan_pca.c:16-68 (norm 2) and cluster_dbscan_kmer.c:18-77:
    returns (stdout text, {file: bytes}, split set); file keys are
    'out-%03d' and, with_values, 'vals-%03d'."""
    acc = vec_acc(values, norm)
    lines: List[str] = []
    s = dbscan(values.shape[1], eps, minpts, region_query_from_acc(acc, norm), lines)
    lines.append("%u clusters obtained \n" % s.n_clusters)
    files = {}
    if s.n_clusters < 500:
        for i, c in enumerate(s.clusters):
            files["out-%03d" % i] = cluster_file_text(c, seqs)
        if with_values:
            for i, c in enumerate(s.clusters):
                files["vals-%03d" % i] = values_file_text(c, values)
    return "".join(lines), files, s


# --------------------------------------------------------------------------
# k-mer frequencies: kmers.c:26-243, 288-389 (fasta2kmer.c)
# --------------------------------------------------------------------------

PROTEIN_ALPHABET = b"ACDEFGHIKLMNPQRSTVWY"  # kmers.c:139-142


def kmer_names(k: int, protein: bool) -> List[bytes]:
    """The k-mer strings in output order: DNA index i spells base p from bits
    2p, 2p+1 of i (A C G T = 0..3, first base in the low bits;
    gen_dna_string_from_binary, kmers.c:26-65); protein k-mers in the
    recursion's order, first residue varying slowest (kmers.c:105-157)."""
    if protein:
        out = [b""]
        for _ in range(k):
            out = [w + bytes([c]) for w in out for c in PROTEIN_ALPHABET]
        return out
    return [bytes(b"ACGT"[(i >> (2 * p)) & 3] for p in range(k)) for i in range(4 ** k)]


def kmer_frequencies(seqs: Sequence[bytes], k: int, protein: bool) -> np.ndarray:
    """frequence_of_kmer_in_sequence (kmers.c:159-243) for every sequence and
    k-mer: occurrences at the start positions i < len - k (the window that
    ends at the last residue is not counted), in a short int that wraps."""
    names = kmer_names(k, protein)
    index = {w: j for j, w in enumerate(names)}
    out = np.zeros((len(seqs), len(names)), np.int64)
    for s, seq in enumerate(seqs):
        for i in range(max(0, len(seq) - k)):
            j = index.get(seq[i:i + k])
            if j is not None:
                out[s, j] += 1
    return ((out + 32768) % 65536 - 32768).astype(np.int16)


def kmer_text(freq: np.ndarray) -> bytes:
    """write_kmer_base (kmers.c:378-389): "sequence_<i>\\t", the counts
    "%hi"-formatted and tab-separated, a line break after the last."""
    lines = []
    for i, row in enumerate(freq.tolist()):
        lines.append("sequence_%i\t" % i + "\t".join(str(v) for v in row) + "\n")
    return "".join(lines).encode()
