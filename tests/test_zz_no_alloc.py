"""GPU test of mnhn_sw_one_vs_all_noalloc, the entry point that allocates
nothing: its device output equals mnhn_sw_one_vs_all's bit for bit, over calls
in a row whose inputs shrink, through the C ABI."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip(lib):
    """The HIP runtime the library is linked against: symbols looked up
    through the library's own handle resolve in its dependencies."""
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return lib


def _set(rng, n, lo, hi, alphabet=b"ACGT"):
    alpha = np.frombuffer(alphabet, np.uint8)
    base = alpha[rng.integers(0, len(alpha), hi)]
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi + 1))
        s = base[:L].copy()
        flip = rng.random(L) < rng.uniform(0.02, 0.6)
        s[flip] = alpha[rng.integers(0, len(alpha), int(flip.sum()))]
        out.append(s.tobytes())
    return out


def test_one_vs_all_noalloc_equals_one_vs_all(pkg):
    """Calls in a row on shrinking inputs -- 3,000 sequences of 60-700
    residues (fp16, int16 and strip-32 passes), then 700 of 40-300, then
    300 over a second alphabet (register-profile kernel), then 7 of
    16,500-18,000 residues (32-bit pass) -- each equal to mnhn_sw_one_vs_all; the
    device buffer is filled with a sentinel first, so every entry is written."""
    lib = pkg.load_library()
    lib.mnhn_sw_one_vs_all_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.mnhn_sw_one_vs_all_noalloc.restype = ctypes.c_int
    hip = _hip(lib)
    rng = np.random.default_rng(2024)
    sets = [_set(rng, 3000, 60, 700), _set(rng, 700, 40, 300), _set(rng, 300, 30, 200, b"ACGTN"),
            _set(rng, 7, 16500, 18000)]
    cap = max(len(s) for s in sets)
    dptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dptr), cap * 4) == 0
    try:
        for seqs in sets:
            n = len(seqs)
            with pkg.Engine(seqs) as eng:
                # the long set: its longest sequence, whose pairs with the other
                # long ones run the 32-bit pass
                qs = sorted({0, n // 2, n - 1}) if n > 7 else [int(np.argmax([len(x) for x in seqs]))]
                for q in qs:
                    print(f"noalloc one-vs-all: n={n} query={q} (length {len(seqs[q])})", flush=True)
                    want = eng.one_vs_all(q)
                    assert hip.hipMemset(dptr, 0xFF, n * 4) == 0
                    rc = lib.mnhn_sw_one_vs_all_noalloc(eng.ctx, q, dptr)
                    assert rc == 0, rc
                    got = np.empty(n, np.int32)
                    assert hip.hipMemcpy(got.ctypes.data, dptr, n * 4, 2) == 0
                    assert np.array_equal(got, want), (n, q, np.flatnonzero(got != want)[:5])
                assert lib.mnhn_sw_one_vs_all_noalloc(eng.ctx, n, dptr) == 1
                assert lib.mnhn_sw_one_vs_all_noalloc(eng.ctx, 0, None) == 1
    finally:
        hip.hipFree(dptr)


def test_one_vs_all_noalloc_rejects_sequences_beyond_its_scratch(pkg):
    """A context whose longest sequence exceeds 65,534 residues does not fit
    one workgroup's workspaces in the static scratch: MNHN_ERR_RANGE, and no
    kernel runs."""
    lib = pkg.load_library()
    lib.mnhn_sw_one_vs_all_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    hip = _hip(lib)
    dptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dptr), 8) == 0
    try:
        seqs = [b"ACGT" * 16384, b"ACGTTGCA" * 10]
        with pkg.Engine(seqs) as eng:
            assert lib.mnhn_sw_one_vs_all_noalloc(eng.ctx, 1, dptr) == 5
    finally:
        hip.hipFree(dptr)
