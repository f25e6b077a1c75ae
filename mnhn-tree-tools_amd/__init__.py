"""Python binding of the gfx950 Smith-Waterman / DBSCAN library (ctypes).

The compute path is the HIP engine in libmnhn_sw.so (built in-tree by
``make -C mnhn-tree-tools_amd``); this module only marshals arguments.  It
has no CPU fallback: if the library is missing or no GPU is usable, calls
raise.

Import it by path (the directory name is not a Python identifier):

    spec = importlib.util.spec_from_file_location(
        "mnhn_tree_tools_amd", "<repo>/mnhn-tree-tools_amd/__init__.py")
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmnhn_sw.so")
BIN_DIR = os.path.join(PKG_DIR, "bin")

MNHN_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "out of memory",
          4: "no usable GPU", 5: "epsilon above the 16-bit list range, or work beyond 32-bit offsets"}

# every entry point declared in include/mnhn_sw.h and include/mnhn_host.h
ABI_SYMBOLS = [
    "mnhn_sw_create", "mnhn_sw_destroy", "mnhn_sw_one_vs_all", "mnhn_sw_pairs",
    "mnhn_sw_matrix", "mnhn_sw_block", "mnhn_sw_eps_csr", "mnhn_sw_eps_count", "mnhn_csr_free",
    "mnhn_sw_size", "mnhn_sw_data_hash", "mnhn_sw_alphabet", "mnhn_sw_last_timing",
    "mnhn_sw_set_tuning", "mnhn_sw_autotune", "mnhn_sw_strerror", "mnhn_score", "mnhn_sw_one_vs_all_noalloc",
]
PROPOSER_SYMBOLS = [
    "mnhn_kmer_profiles", "mnhn_encoder_create", "mnhn_encoder_destroy",
    "mnhn_encoder_train", "mnhn_encoder_embed", "mnhn_encoder_set_margin",
    "mnhn_encoder_set_margin_table", "mnhn_encoder_set_cosine_table", "mnhn_encoder_margin", "mnhn_encoder_save", "mnhn_encoder_load",
    "mnhn_prefilter_eps_csr",
]
VEC_SYMBOLS = [
    "mnhn_vec_create", "mnhn_vec_destroy", "mnhn_vec_eps_csr", "mnhn_vcsr_free", "mnhn_vec_last_timing",
    "mnhn_kmer_columns", "mnhn_kmer_frequencies",
]
HOST_SYMBOLS = [
    "dataset_from_fasta", "free_sequences_from_dataset", "store_split_set",
    "read_split_set", "create_single_cluster_file", "create_cluster_files",
    "data_not_in_clusters", "intersection_of_clusters",
    "generate_split_set_relation", "free_split_set_and_associated_clusters",
    "mnhn_gpu_initialization", "mnhn_gpu_destroy", "mnhn_gpu_set_lookahead",
    "dbscan_SW_GPU", "dbscan_SW", "mnhn_dbscan_csr", "mnhn_csr_save", "mnhn_csr_load", "adaptive_dbscan",
    "dataset_to_fasta", "filter_ds_by_size", "filter_by_SW_distance",
    "smith_waterman_distances_matrix", "mean_from_smith_waterman_datasets",
    "sigma_from_smith_waterman_datasets", "silhouette_from_smith_waterman_datasets",
    "mean_from_smith_waterman_distance_matrix", "sigma_from_smith_waterman_distance_matrix",
    "shortest_longest_distance_in_matrix", "load_projections_from_file_into_dataset",
    "free_values_from_dataset", "print_SW_PCA_L_comparison",
    "mnhn_gpu_initialization_values", "dbscan_L1_GPU", "dbscan_L2_GPU", "dbscan_L1", "dbscan_L2",
    "mnhn_dbscan_vcsr", "mnhn_dbscan_csr_sweep", "mnhn_dbscan_vcsr_sweep", "shape_from_kmer_file", "load_kmer_from_file_into_dataset",
    "create_single_cluster_file_with_values", "create_cluster_files_with_values", "free_dataset",
    "number_of_kmers_from_length", "number_of_protein_kmers_from_length", "frequencies_from_dataset",
    "write_kmer_base", "free_kmer_frequencies",
]


class MnhnError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("match", ctypes.c_int32), ("mismatch", ctypes.c_int32),
                ("gap", ctypes.c_int32)]


class Csr(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int32), ("eps_max", ctypes.c_int32),
                ("nnz", ctypes.c_int64),
                ("row_ptr", ctypes.POINTER(ctypes.c_int64)),
                ("col", ctypes.POINTER(ctypes.c_int32)),
                ("dist", ctypes.POINTER(ctypes.c_uint16))]


class Vcsr(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int32), ("norm", ctypes.c_int32), ("thr", ctypes.c_float),
                ("reserved", ctypes.c_int32), ("nnz", ctypes.c_int64),
                ("row_ptr", ctypes.POINTER(ctypes.c_int64)), ("col", ctypes.POINTER(ctypes.c_int32)),
                ("val", ctypes.POINTER(ctypes.c_float))]


NORM_L1, NORM_L2 = 1, 2


class Timing(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("cells", ctypes.c_double), ("pairs", ctypes.c_double),
                ("lane_bits", ctypes.c_int32), ("kernel", ctypes.c_int32),
                ("strip", ctypes.c_int32),
                ("waves", ctypes.c_int32), ("items", ctypes.c_int32), ("mixed_launch", ctypes.c_int32),
                ("form_cells", ctypes.c_double * 3), ("form_ms", ctypes.c_double * 3)]


class Tuning(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("strip", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("gcups", ctypes.c_double),
                ("from_cache", ctypes.c_int32)]


class TrainParams(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
                ("eps_train", ctypes.c_int32), ("neighbor_weight", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("weight_decay", ctypes.c_float),
                ("lr_decay", ctypes.c_int32), ("far_weight", ctypes.c_float),
                ("far_gap", ctypes.c_float)]


class TrainStats(ctypes.Structure):
    _fields_ = [("first_loss", ctypes.c_double), ("final_loss", ctypes.c_double),
                ("seconds", ctypes.c_double), ("steps", ctypes.c_int64)]


class PrefilterStats(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int64), ("neighbours", ctypes.c_int64),
                ("profile_ms", ctypes.c_double), ("embed_ms", ctypes.c_double),
                ("propose_ms", ctypes.c_double), ("verify_ms", ctypes.c_double),
                ("csr_ms", ctypes.c_double), ("tiles", ctypes.c_int64), ("cells", ctypes.c_double),
                ("verify_kernel_ms", ctypes.c_double), ("stage2_tiles", ctypes.c_int64),
                ("propose_kernel_ms", ctypes.c_double), ("counts_int8", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


_LIB = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libmnhn_sw.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise MnhnError(f"{path} not built: run `make -C {PKG_DIR}`")
    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.mnhn_sw_create.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(Params),
                                   ctypes.POINTER(vp)]
    lib.mnhn_sw_destroy.argtypes = [vp]
    lib.mnhn_sw_destroy.restype = None
    lib.mnhn_sw_one_vs_all.argtypes = [vp, ctypes.c_int, vp]
    lib.mnhn_sw_pairs.argtypes = [vp, vp, i64, vp]
    lib.mnhn_sw_matrix.argtypes = [vp, vp]
    lib.mnhn_sw_eps_csr.argtypes = [vp, i32, ctypes.POINTER(Csr)]
    lib.mnhn_sw_eps_count.argtypes = [vp, i32, ctypes.POINTER(i64)]
    lib.mnhn_csr_free.argtypes = [ctypes.POINTER(Csr)]
    lib.mnhn_csr_free.restype = None
    lib.mnhn_sw_size.argtypes = [vp]
    lib.mnhn_sw_block.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, vp]
    lib.mnhn_sw_data_hash.argtypes = [vp]
    lib.mnhn_sw_data_hash.restype = ctypes.c_uint64
    lib.mnhn_csr_save.argtypes = [ctypes.POINTER(Csr), ctypes.c_uint64, ctypes.c_char_p]
    lib.mnhn_csr_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(Csr)]
    lib.mnhn_sw_alphabet.argtypes = [vp]
    lib.mnhn_sw_last_timing.argtypes = [vp, ctypes.POINTER(Timing)]
    lib.mnhn_sw_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.mnhn_sw_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Tuning)]
    lib.mnhn_sw_strerror.argtypes = [ctypes.c_int]
    lib.mnhn_sw_strerror.restype = ctypes.c_char_p
    lib.mnhn_kmer_profiles.argtypes = [vp, vp]
    lib.mnhn_encoder_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(vp)]
    lib.mnhn_encoder_destroy.argtypes = [vp]
    lib.mnhn_encoder_destroy.restype = None
    lib.mnhn_encoder_train.argtypes = [vp, vp, vp, vp, i64, ctypes.POINTER(TrainParams),
                                       ctypes.POINTER(TrainStats)]
    lib.mnhn_encoder_embed.argtypes = [vp, vp, vp]
    lib.mnhn_encoder_set_margin.argtypes = [vp, ctypes.c_float]
    lib.mnhn_encoder_set_margin_table.argtypes = [vp, vp, vp, ctypes.c_int]
    lib.mnhn_encoder_set_cosine_table.argtypes = [vp, vp, vp, ctypes.c_int]
    lib.mnhn_encoder_margin.argtypes = [vp]
    lib.mnhn_encoder_margin.restype = ctypes.c_float
    lib.mnhn_encoder_save.argtypes = [vp, ctypes.c_char_p]
    lib.mnhn_encoder_load.argtypes = [vp, ctypes.c_char_p]
    lib.mnhn_prefilter_eps_csr.argtypes = [vp, vp, i32, ctypes.POINTER(Csr),
                                           ctypes.POINTER(PrefilterStats)]
    lib.mnhn_vec_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    lib.mnhn_vec_destroy.argtypes = [vp]
    lib.mnhn_vec_destroy.restype = None
    lib.mnhn_vec_eps_csr.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.POINTER(Vcsr)]
    lib.mnhn_vcsr_free.argtypes = [ctypes.POINTER(Vcsr)]
    lib.mnhn_vcsr_free.restype = None
    lib.mnhn_vec_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
    lib.mnhn_kmer_columns.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.mnhn_kmer_columns.restype = ctypes.c_size_t
    lib.mnhn_kmer_frequencies.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    _LIB = lib
    return lib


def _check(rc: int, what: str) -> None:
    if rc != MNHN_OK:
        raise MnhnError(f"{what}: {ERRORS.get(rc, rc)} (code {rc})")


class Engine:
    """One uploaded sequence set (mnhn_sw_ctx)."""

    def __init__(self, seqs: Sequence[bytes], params: Optional[Tuple[int, int, int]] = None):
        self.lib = load_library()
        self.n = len(seqs)
        self.lens = np.array([len(s) for s in seqs], dtype=np.uint64)
        bufs = [ctypes.create_string_buffer(bytes(s), len(s) + 1) for s in seqs]
        ptrs = (ctypes.c_char_p * max(self.n, 1))(
            *[ctypes.cast(b, ctypes.c_char_p) for b in bufs])
        lens = (ctypes.c_size_t * max(self.n, 1))(*[len(s) for s in seqs])
        p = Params(*params) if params is not None else None
        ctx = ctypes.c_void_p()
        rc = self.lib.mnhn_sw_create(ptrs, lens, self.n,
                                     ctypes.byref(p) if p is not None else None,
                                     ctypes.byref(ctx))
        _check(rc, "mnhn_sw_create")
        self.ctx = ctx

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.mnhn_sw_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def alphabet(self) -> int:
        return int(self.lib.mnhn_sw_alphabet(self.ctx))

    def set_tuning(self, kernel: int = 0, strip: int = 0, waves_per_cu: int = 0) -> None:
        """kernel: 0 auto, 1 register profile, 2 LDS profile; strip 16/32/64."""
        _check(self.lib.mnhn_sw_set_tuning(self.ctx, kernel, strip, waves_per_cu), "set_tuning")

    def autotune(self, use_cache: bool = True) -> dict:
        """Time the candidate kernel settings on a sample of this set's
        all-pairs work (or reuse the cached choice) and keep the fastest."""
        t = Tuning()
        _check(self.lib.mnhn_sw_autotune(self.ctx, 1 if use_cache else 0, ctypes.byref(t)),
               "mnhn_sw_autotune")
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def one_vs_all(self, query: int) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(self.lib.mnhn_sw_one_vs_all(self.ctx, int(query), out.ctypes.data),
               "mnhn_sw_one_vs_all")
        return out

    def pairs(self, ij: np.ndarray) -> np.ndarray:
        ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1, 2)
        out = np.empty(len(ij), dtype=np.int32)
        _check(self.lib.mnhn_sw_pairs(self.ctx, ij.ctypes.data, len(ij), out.ctypes.data),
               "mnhn_sw_pairs")
        return out

    def matrix(self) -> np.ndarray:
        out = np.empty((self.n, self.n), dtype=np.int32)
        _check(self.lib.mnhn_sw_matrix(self.ctx, out.ctypes.data), "mnhn_sw_matrix")
        return out

    def block(self, rows, cols) -> np.ndarray:
        """int32 [len(rows), len(cols)] distances d(seq[rows[r]], seq[cols[k]])."""
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.empty((len(rows), len(cols)), dtype=np.int32)
        _check(self.lib.mnhn_sw_block(self.ctx, rows.ctypes.data, len(rows), cols.ctypes.data, len(cols),
                                      out.ctypes.data), "mnhn_sw_block")
        return out

    @property
    def data_hash(self) -> int:
        """FNV-1a of scores, count, lengths and bytes (the .ecsr file key)."""
        return int(self.lib.mnhn_sw_data_hash(self.ctx))

    def eps_csr(self, eps_max: int):
        """Returns (row_ptr int64[n+1], col int32[nnz], dist uint16[nnz])."""
        csr = Csr()
        _check(self.lib.mnhn_sw_eps_csr(self.ctx, int(eps_max), ctypes.byref(csr)),
               "mnhn_sw_eps_csr")
        return _csr_to_numpy(self.lib, csr, self.n)

    def eps_count(self, eps_max: int) -> int:
        v = ctypes.c_int64(0)
        _check(self.lib.mnhn_sw_eps_count(self.ctx, int(eps_max), ctypes.byref(v)),
               "mnhn_sw_eps_count")
        return int(v.value)

    def last_timing(self) -> dict:
        t = Timing()
        _check(self.lib.mnhn_sw_last_timing(self.ctx, ctypes.byref(t)), "last_timing")
        out = {f: getattr(t, f) for f, _ in Timing._fields_ if f != "reserved"}
        for f in ("form_cells", "form_ms"):
            out[f] = list(out[f])
        return out

    def kmer_profiles(self) -> np.ndarray:
        """5-mer profiles (n x 1024 fp32, L1-normalised) computed on the GPU."""
        out = np.empty((self.n, 1024), dtype=np.float32)
        _check(self.lib.mnhn_kmer_profiles(self.ctx, out.ctypes.data), "mnhn_kmer_profiles")
        return out


class _CsrOwner:
    """Keeps a library-owned CSR alive; frees it with mnhn_csr_free when the
    last array viewing its buffers is gone."""

    def __init__(self, lib, csr):
        self.lib, self.csr = lib, csr

    def __del__(self):
        try:
            self.lib.mnhn_csr_free(ctypes.byref(self.csr))
        except Exception:
            pass


class _OwnedArray(np.ndarray):
    """ndarray view over library memory that holds the memory's owner."""


def _view(ptr, n, owner):
    a = np.ctypeslib.as_array(ptr, shape=(n,)).view(_OwnedArray)
    a._owner = owner
    return a


def _csr_to_numpy(lib, csr, n):
    """(row_ptr, col, dist) as zero-copy views of the library's CSR buffers."""
    owner = _CsrOwner(lib, csr)
    nnz = int(csr.nnz)
    rp = _view(csr.row_ptr, n + 1, owner)
    col = _view(csr.col, nnz, owner) if nnz else np.zeros(0, np.int32)
    dist = _view(csr.dist, nnz, owner) if nnz else np.zeros(0, np.uint16)
    return rp, col, dist


def vec_threshold(norm: int, eps: float) -> np.float32:
    """The value the accumulator is compared with: eps * eps in float32 for
    L2 (dbscan.c:577), eps for L1."""
    e = np.float32(eps)
    return np.float32(e * e) if norm == NORM_L2 else e


class _VcsrOwner:
    def __init__(self, lib, csr):
        self.lib, self.csr = lib, csr

    def __del__(self):
        try:
            self.lib.mnhn_vcsr_free(ctypes.byref(self.csr))
        except Exception:
            pass


class VecEngine:
    """Epsilon lists of float vectors under the reference's L1 / L2 scans on
    the GPU (mnhn_vec.h).  points: (n, dims) float32 array."""

    def __init__(self, points: np.ndarray):
        self.lib = load_library()
        pts = np.asarray(points, dtype=np.float32)
        if pts.ndim != 2:
            raise MnhnError("points must be (n, dims)")
        self.n, self.dims = pts.shape
        cols = [np.ascontiguousarray(pts[:, k]) for k in range(self.dims)]
        ptrs = (ctypes.c_void_p * max(self.dims, 1))(*[c.ctypes.data for c in cols])
        self.ctx = ctypes.c_void_p()
        _check(self.lib.mnhn_vec_create(ptrs, self.dims, self.n, ctypes.byref(self.ctx)),
               "mnhn_vec_create")

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.mnhn_vec_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def eps_csr(self, norm: int, thr: float):
        """(row_ptr int64[n+1], col int32[nnz], val float32[nnz]): per row every
        point (itself included) whose accumulator is <= thr, ascending."""
        csr = Vcsr()
        _check(self.lib.mnhn_vec_eps_csr(self.ctx, int(norm), ctypes.c_float(thr), ctypes.byref(csr)),
               "mnhn_vec_eps_csr")
        owner = _VcsrOwner(self.lib, csr)
        nnz = int(csr.nnz)
        rp = _view(csr.row_ptr, self.n + 1, owner)
        col = _view(csr.col, nnz, owner) if nnz else np.zeros(0, np.int32)
        val = _view(csr.val, nnz, owner) if nnz else np.zeros(0, np.float32)
        return rp, col, val

    def last_timing(self) -> dict:
        ms, pairs = ctypes.c_double(0), ctypes.c_double(0)
        _check(self.lib.mnhn_vec_last_timing(self.ctx, ctypes.byref(ms), ctypes.byref(pairs)),
               "mnhn_vec_last_timing")
        return {"kernel_ms": ms.value, "pairs": pairs.value}


def kmer_frequencies(seqs: Sequence[bytes], k: int, protein: bool = False) -> np.ndarray:
    """(n, 4^k or 20^k) int16 k-mer counts on the GPU (mnhn_kmer.h): the
    table of the reference's frequencies_from_dataset."""
    lib = load_library()
    nk = int(lib.mnhn_kmer_columns(int(k), int(bool(protein))))
    if nk == 0:
        raise MnhnError(f"k-mer length {k} not supported")
    n = len(seqs)
    bufs = [ctypes.create_string_buffer(x, len(x) + 1) for x in seqs]
    ptrs = (ctypes.c_char_p * max(n, 1))(*[ctypes.cast(b, ctypes.c_char_p) for b in bufs])
    lens = (ctypes.c_size_t * max(n, 1))(*[len(x) for x in seqs])
    out = np.empty((n, nk), np.int16)
    _check(lib.mnhn_kmer_frequencies(ptrs, lens, n, int(k), int(bool(protein)), out.ctypes.data),
           "mnhn_kmer_frequencies")
    return out


class Encoder:
    """Learned candidate proposer (mnhn_encoder): 5-mer profile -> 128-d MLP."""

    def __init__(self, seed: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        _check(self.lib.mnhn_encoder_create(ctypes.c_uint64(seed), ctypes.byref(h)),
               "mnhn_encoder_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.mnhn_encoder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def train(self, eng: Engine, ij: np.ndarray, dist: np.ndarray, epochs: int = 4,
              batch: int = 4096, lr: float = 1e-3, eps_train: int = 400,
              neighbor_weight: float = 4.0, seed: int = 0, weight_decay: float = 0.0,
              lr_decay: bool = False, far_weight: float = 0.0, far_gap: float = 0.0) -> dict:
        ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1, 2)
        dist = np.ascontiguousarray(dist, dtype=np.int32)
        p = TrainParams(epochs, batch, lr, eps_train, neighbor_weight, seed, weight_decay,
                        int(bool(lr_decay)), far_weight, far_gap)
        st = TrainStats()
        _check(self.lib.mnhn_encoder_train(self.h, eng.ctx, ij.ctypes.data, dist.ctypes.data,
                                           len(ij), ctypes.byref(p), ctypes.byref(st)),
               "mnhn_encoder_train")
        return {f: getattr(st, f) for f, _ in TrainStats._fields_}

    def embed(self, eng: Engine) -> np.ndarray:
        out = np.empty((eng.n, 128), dtype=np.float32)
        _check(self.lib.mnhn_encoder_embed(self.h, eng.ctx, out.ctypes.data), "mnhn_encoder_embed")
        return out

    @property
    def margin(self) -> float:
        return float(self.lib.mnhn_encoder_margin(self.h))

    @margin.setter
    def margin(self, v: float) -> None:
        _check(self.lib.mnhn_encoder_set_margin(self.h, float(v)), "mnhn_encoder_set_margin")

    def set_margin_table(self, max_len, margins) -> None:
        """Margin per length bucket of the pair's longer sequence (first bucket
        with length <= max_len[b]; the last margin above the last bound)."""
        ml = np.ascontiguousarray(max_len, dtype=np.int32)
        mg = np.ascontiguousarray(margins, dtype=np.float32)
        if ml.shape != mg.shape:
            raise MnhnError("max_len and margins differ in length")
        _check(self.lib.mnhn_encoder_set_margin_table(self.h, ml.ctypes.data, mg.ctypes.data, len(ml)),
               "mnhn_encoder_set_margin_table")

    def set_cosine_table(self, max_len, cos_min) -> None:
        """Minimum raw 5-mer profile cosine per length bucket of the pair's
        longer sequence (negative: no test in that bucket)."""
        ml = np.ascontiguousarray(max_len, dtype=np.int32)
        cm = np.ascontiguousarray(cos_min, dtype=np.float32)
        if ml.shape != cm.shape:
            raise MnhnError("max_len and cos_min differ in length")
        _check(self.lib.mnhn_encoder_set_cosine_table(self.h, ml.ctypes.data, cm.ctypes.data, len(ml)),
               "mnhn_encoder_set_cosine_table")

    def save(self, path: str) -> None:
        _check(self.lib.mnhn_encoder_save(self.h, path.encode()), "mnhn_encoder_save")

    def load(self, path: str) -> None:
        _check(self.lib.mnhn_encoder_load(self.h, path.encode()), "mnhn_encoder_load")

    def prefilter_csr(self, eng: Engine, eps_max: int):
        """(row_ptr, col, dist, stats): exact distances of the proposed pairs with d <= eps_max."""
        csr = Csr()
        st = PrefilterStats()
        _check(self.lib.mnhn_prefilter_eps_csr(self.h, eng.ctx, int(eps_max), ctypes.byref(csr),
                                               ctypes.byref(st)), "mnhn_prefilter_eps_csr")
        rp, col, dist = _csr_to_numpy(self.lib, csr, eng.n)
        return rp, col, dist, {f: getattr(st, f) for f, _ in PrefilterStats._fields_}


def load_encoder_weights(path: str):
    """Read a file written by Encoder.save: ({"margin": (max_len[], margin[]),
    "cosine": (max_len[], cos_min[])}, [(Wt [out][in], b [out]) x 3])."""
    def table(f):
        nb = int(np.frombuffer(f.read(4), "<i4")[0])
        return (np.frombuffer(f.read(4 * nb), "<i4").copy(), np.frombuffer(f.read(4 * nb), "<f4").copy())

    off = (np.array([2**31 - 1], np.int32), np.array([-1.0], np.float32))
    with open(path, "rb") as f:
        magic = f.read(8)
        if magic == b"MNHNENC1":
            tables = {"margin": (np.array([2**31 - 1], np.int32), np.frombuffer(f.read(4), "<f4").copy()),
                      "cosine": off}
        elif magic == b"MNHNENC2":
            tables = {"margin": table(f), "cosine": off}
        elif magic == b"MNHNENC3":
            m = table(f)
            tables = {"margin": m, "cosine": table(f)}
        else:
            raise MnhnError(f"{path}: not an encoder file")
        layers = []
        for _ in range(3):
            din, dout = np.frombuffer(f.read(8), "<i4")
            w = np.frombuffer(f.read(4 * din * dout), "<f4").reshape(dout, din)
            b = np.frombuffer(f.read(4 * dout), "<f4")
            layers.append((w, b))
    return tables, layers


class _Cluster(ctypes.Structure):
    _fields_ = [("members", ctypes.POINTER(ctypes.c_int)), ("n_members", ctypes.c_int),
                ("id", ctypes.c_int)]


class _SplitSet(ctypes.Structure):
    _fields_ = [("n_clusters", ctypes.c_int), ("clusters", ctypes.POINTER(_Cluster))]


def dbscan_csr(rp: np.ndarray, col: np.ndarray, dist: np.ndarray, eps_max: int, eps: float,
               minpts: int):
    """One DBSCAN pass (the library's C host code) over epsilon lists; returns
    the clusters as lists of member indices in the reference's order."""
    lib = load_library()
    lib.mnhn_dbscan_csr.restype = _SplitSet
    lib.mnhn_dbscan_csr.argtypes = [ctypes.POINTER(Csr), ctypes.c_float, ctypes.c_int]
    lib.free_split_set_and_associated_clusters.argtypes = [_SplitSet]
    lib.free_split_set_and_associated_clusters.restype = None
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    dist = np.ascontiguousarray(dist, dtype=np.uint16)
    csr = Csr(len(rp) - 1, int(eps_max), len(col),
              rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
              col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
              dist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    s = lib.mnhn_dbscan_csr(ctypes.byref(csr), float(eps), int(minpts))
    try:
        out = []
        for k in range(s.n_clusters):
            c = s.clusters[k]
            out.append(np.ctypeslib.as_array(c.members, shape=(c.n_members,)).tolist()
                       if c.n_members else [])
    finally:
        lib.free_split_set_and_associated_clusters(s)
    return out


def dbscan_vcsr(rp: np.ndarray, col: np.ndarray, val: np.ndarray, norm: int, thr: float, eps: float,
                minpts: int):
    """One L1 / L2 DBSCAN pass (the library's C host code) over vector
    epsilon lists computed up to thr; returns the clusters' member lists."""
    lib = load_library()
    lib.mnhn_dbscan_vcsr.restype = _SplitSet
    lib.mnhn_dbscan_vcsr.argtypes = [ctypes.POINTER(Vcsr), ctypes.c_float, ctypes.c_int]
    lib.free_split_set_and_associated_clusters.argtypes = [_SplitSet]
    lib.free_split_set_and_associated_clusters.restype = None
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float32)
    csr = Vcsr(len(rp) - 1, int(norm), float(thr), 0, len(col),
               rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
               col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
               val.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    s = lib.mnhn_dbscan_vcsr(ctypes.byref(csr), float(eps), int(minpts))
    try:
        out = []
        for k in range(s.n_clusters):
            c = s.clusters[k]
            out.append(np.ctypeslib.as_array(c.members, shape=(c.n_members,)).tolist()
                       if c.n_members else [])
    finally:
        lib.free_split_set_and_associated_clusters(s)
    return out


def _sweep(fn, csr, eps, minpts, symmetric):
    lib = load_library()
    eps = np.ascontiguousarray(eps, dtype=np.float32)
    sets = (_SplitSet * max(len(eps), 1))()
    f = getattr(lib, fn)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(type(csr)), ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.POINTER(_SplitSet)]
    lib.free_split_set_and_associated_clusters.argtypes = [_SplitSet]
    lib.free_split_set_and_associated_clusters.restype = None
    _check(f(ctypes.byref(csr), eps.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(eps), int(minpts),
             int(bool(symmetric)), sets), fn)
    out = []
    for k in range(len(eps)):
        s = sets[k]
        try:
            out.append([np.ctypeslib.as_array(s.clusters[c].members, shape=(s.clusters[c].n_members,)).tolist()
                        if s.clusters[c].n_members else [] for c in range(s.n_clusters)])
        finally:
            lib.free_split_set_and_associated_clusters(s)
    return out


def dbscan_csr_sweep(rp: np.ndarray, col: np.ndarray, dist: np.ndarray, eps_max: int, eps, minpts: int,
                     symmetric: bool = True):
    """DBSCAN passes at every epsilon of eps over the same Smith-Waterman
    lists (mnhn_dbscan_csr_sweep: the lists ordered by distance once); one
    list of clusters per epsilon, each as dbscan_csr returns it.  symmetric:
    every (p, q, d) is listed as (q, p, d) too, as in the engine's lists."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    dist = np.ascontiguousarray(dist, dtype=np.uint16)
    csr = Csr(len(rp) - 1, int(eps_max), len(col), rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
              col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
              dist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    return _sweep("mnhn_dbscan_csr_sweep", csr, eps, minpts, symmetric)


def dbscan_vcsr_sweep(rp: np.ndarray, col: np.ndarray, val: np.ndarray, norm: int, thr: float, eps,
                      minpts: int, symmetric: bool = True):
    """L1 / L2 DBSCAN passes at every epsilon of eps over the same vector
    lists (mnhn_dbscan_vcsr_sweep); one list of clusters per epsilon."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float32)
    csr = Vcsr(len(rp) - 1, int(norm), float(thr), 0, len(col),
               rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
               col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
               val.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return _sweep("mnhn_dbscan_vcsr_sweep", csr, eps, minpts, symmetric)


def csr_save(path: str, rp: np.ndarray, col: np.ndarray, dist: np.ndarray, eps_max: int,
             data_hash: int) -> None:
    """Write epsilon lists to an .ecsr file (mnhn_csr_save)."""
    lib = load_library()
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    dist = np.ascontiguousarray(dist, dtype=np.uint16)
    csr = Csr(len(rp) - 1, int(eps_max), len(col),
              rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
              col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
              dist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    _check(lib.mnhn_csr_save(ctypes.byref(csr), ctypes.c_uint64(data_hash), path.encode()),
           "mnhn_csr_save")


def csr_load(path: str):
    """Read an .ecsr file: (row_ptr, col, dist, eps_max, data_hash)."""
    lib = load_library()
    csr = Csr()
    h = ctypes.c_uint64(0)
    _check(lib.mnhn_csr_load(path.encode(), ctypes.byref(h), ctypes.byref(csr)), "mnhn_csr_load")
    eps_max = int(csr.eps_max)
    rp, col, dist = _csr_to_numpy(lib, csr, int(csr.n_rows))
    return rp, col, dist, eps_max, int(h.value)

