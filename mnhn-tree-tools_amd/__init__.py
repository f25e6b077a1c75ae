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
          4: "no usable GPU", 5: "score range exceeds the 16-bit lane format"}

# every entry point declared in include/mnhn_sw.h and include/mnhn_host.h
ABI_SYMBOLS = [
    "mnhn_sw_create", "mnhn_sw_destroy", "mnhn_sw_one_vs_all", "mnhn_sw_pairs",
    "mnhn_sw_matrix", "mnhn_sw_eps_csr", "mnhn_sw_eps_count", "mnhn_csr_free",
    "mnhn_sw_size", "mnhn_sw_alphabet", "mnhn_sw_last_timing",
    "mnhn_sw_set_tuning", "mnhn_sw_autotune", "mnhn_sw_strerror",
]
HOST_SYMBOLS = [
    "dataset_from_fasta", "free_sequences_from_dataset", "store_split_set",
    "read_split_set", "create_single_cluster_file", "create_cluster_files",
    "data_not_in_clusters", "intersection_of_clusters",
    "generate_split_set_relation", "free_split_set_and_associated_clusters",
    "mnhn_gpu_initialization", "mnhn_gpu_destroy", "mnhn_gpu_set_lookahead",
    "dbscan_SW_GPU", "adaptive_dbscan",
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
                ("dist", ctypes.POINTER(ctypes.c_int16))]


class Timing(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("cells", ctypes.c_double), ("pairs", ctypes.c_double),
                ("lane_bits", ctypes.c_int32), ("kernel", ctypes.c_int32),
                ("strip", ctypes.c_int32),
                ("waves", ctypes.c_int32), ("items", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int32), ("strip", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("gcups", ctypes.c_double),
                ("from_cache", ctypes.c_int32)]


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
    lib.mnhn_sw_alphabet.argtypes = [vp]
    lib.mnhn_sw_last_timing.argtypes = [vp, ctypes.POINTER(Timing)]
    lib.mnhn_sw_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.mnhn_sw_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Tuning)]
    lib.mnhn_sw_strerror.argtypes = [ctypes.c_int]
    lib.mnhn_sw_strerror.restype = ctypes.c_char_p
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

    def eps_csr(self, eps_max: int):
        """Returns (row_ptr int64[n+1], col int32[nnz], dist int16[nnz])."""
        csr = Csr()
        _check(self.lib.mnhn_sw_eps_csr(self.ctx, int(eps_max), ctypes.byref(csr)),
               "mnhn_sw_eps_csr")
        try:
            nnz = int(csr.nnz)
            rp = np.ctypeslib.as_array(csr.row_ptr, shape=(self.n + 1,)).copy()
            col = (np.ctypeslib.as_array(csr.col, shape=(nnz,)).copy() if nnz
                   else np.zeros(0, np.int32))
            dist = (np.ctypeslib.as_array(csr.dist, shape=(nnz,)).copy() if nnz
                    else np.zeros(0, np.int16))
        finally:
            self.lib.mnhn_csr_free(ctypes.byref(csr))
        return rp, col, dist

    def eps_count(self, eps_max: int) -> int:
        v = ctypes.c_int64(0)
        _check(self.lib.mnhn_sw_eps_count(self.ctx, int(eps_max), ctypes.byref(v)),
               "mnhn_sw_eps_count")
        return int(v.value)

    def last_timing(self) -> dict:
        t = Timing()
        _check(self.lib.mnhn_sw_last_timing(self.ctx, ctypes.byref(t)), "last_timing")
        return {f: getattr(t, f) for f, _ in Timing._fields_}
