"""GPU parity tests of the L1 / L2 vector engine (mnhn_vec.h) and the vector
DBSCAN tools: epsilon lists bit-identical to the oracle's accumulators, the
drop-in CLIs byte-identical to the compiled reference on every fixture, and
size-independent properties at larger sizes."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR, golden_arg, vector_cases

pytestmark = pytest.mark.gpu

VEC_CLI_CASES = [c for c in vector_cases() if c[1]["program"] != "fasta2kmer"]


def _expected_lists(acc, thr):
    keep = acc <= np.float32(thr)
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int64)
    return rp, np.nonzero(keep)[1].astype(np.int32), acc[keep].astype(np.float32)


def _check_exact(pkg, oracle, points, norm, thr):
    """Lists of every row against the oracle: same columns, same accumulator
    bits."""
    n = points.shape[0]
    print(f"vec lists n={n} dims={points.shape[1]} norm={norm} thr={thr}", flush=True)
    with pkg.VecEngine(points) as eng:
        rp, col, val = eng.eps_csr(norm, thr)
    acc = oracle.vec_acc(points.T, norm)
    wrp, wcol, wval = _expected_lists(acc, thr)
    assert np.array_equal(rp, wrp)
    assert np.array_equal(col, wcol)
    assert np.array_equal(val.view(np.uint32), wval.view(np.uint32))
    return len(col)


def _thr_for_density(oracle, points, norm, frac):
    acc = oracle.vec_acc(points.T, norm, rows=np.arange(min(64, len(points))))
    return float(np.quantile(acc[np.isfinite(acc)], frac)) if np.isfinite(acc).any() else 1.0


@pytest.mark.parametrize("norm", [1, 2])
@pytest.mark.parametrize("n,dims", [(1, 1), (2, 3), (127, 2), (128, 5), (129, 17), (300, 64), (700, 300),
                                    (1000, 3)])
def test_vec_lists_match_oracle(pkg, oracle, n, dims, norm):
    rng = np.random.default_rng(n * 1000 + dims * 10 + norm)
    pts = (rng.normal(size=(n, dims)) * rng.uniform(0.1, 10, size=dims)).astype(np.float32)
    for frac in (0.05, 0.3):
        _check_exact(pkg, oracle, pts, norm, _thr_for_density(oracle, pts, norm, frac))


@pytest.mark.parametrize("norm", [1, 2])
def test_vec_special_values(pkg, oracle, norm):
    """NaN / inf / denormal / signed-zero coordinates, integer k-mer counts,
    and the thresholds at the edges: negative, zero, +inf."""
    rng = np.random.default_rng(17 + norm)
    pts = rng.integers(0, 6, size=(400, 64)).astype(np.float32)
    pts[3, 5] = np.nan
    pts[9, 0] = np.inf
    pts[10, 0] = -np.inf
    pts[11, 7] = np.float32(1e-41)  # denormal
    pts[12, 7] = np.float32(-1e-41)
    pts[13, :] = -0.0
    pts[14, :] = 0.0
    pts[200:210] = pts[100]  # duplicates: accumulator 0
    for thr in (-1.0, 0.0, 5.0, 60.0, float(np.inf)):
        _check_exact(pkg, oracle, pts, norm, thr)


def test_vec_capacity_retry(pkg, oracle):
    """More pairs than the first pair buffer holds (n * 32 or 2^20): the
    kernel counts past the end and runs again with the exact size."""
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(2100, 4)).astype(np.float32)
    nnz = _check_exact(pkg, oracle, pts, 2, float(np.inf))
    assert nnz == 2100 * 2100


def test_vec_dbscan_matches_oracle(pkg, oracle):
    """Lists from the GPU at a threshold above several epsilons; the host
    DBSCAN over them gives the oracle's clusters for each epsilon."""
    values = oracle.read_kmers(os.path.join(GOLDEN, "inputs", "c2_small_k4.kmer"))
    for norm, e_max, passes in ((2, 20.0, ((10, 4), (14, 4), (20, 3))), (1, 120.0, ((60, 5), (90, 4)))):
        with pkg.VecEngine(values.T) as eng:
            t_max = pkg.vec_threshold(norm, e_max)
            rp, col, val = eng.eps_csr(norm, t_max)
        acc = oracle.vec_acc(values, norm)
        for eps, minpts in passes:
            got = pkg.dbscan_vcsr(rp, col, val, norm, t_max, eps, minpts)
            want = oracle.dbscan(values.shape[1], eps, minpts, oracle.region_query_from_acc(acc, norm))
            assert got == [c.members for c in want.clusters], (norm, eps, minpts)


@pytest.mark.parametrize("n,dims,norm", [(60000, 5, 2), (12000, 1024, 1)])
def test_vec_properties_at_size(pkg, oracle, n, dims, norm):
    """Larger inputs (clustered projections; k-mer-like counts): rows sorted,
    lists symmetric with equal accumulators both ways, every value under the
    threshold, and 48 sampled rows bit-identical to the oracle's."""
    rng = np.random.default_rng(n + dims)
    if dims <= 8:
        centres = rng.normal(scale=4.0, size=(200, dims))
        pts = (centres[rng.integers(0, 200, n)] + rng.normal(scale=0.3, size=(n, dims))).astype(np.float32)
    else:
        pts = rng.poisson(0.4, size=(n, dims)).astype(np.float32)
        pts[n // 2:] = pts[: n - n // 2] + rng.integers(0, 2, size=(n - n // 2, dims))
    thr = _thr_for_density(oracle, pts, norm, 0.01)
    print(f"vec properties n={n} dims={dims} norm={norm} thr={thr}", flush=True)
    with pkg.VecEngine(pts) as eng:
        rp, col, val = eng.eps_csr(norm, thr)
        t = eng.last_timing()
    assert t["pairs"] == n * (n + 1) / 2 and t["kernel_ms"] > 0
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(val <= np.float32(thr))
    key = rows.astype(np.int64) * n + col
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates
    tkey = col.astype(np.int64) * n + rows
    order = np.argsort(tkey)
    assert np.array_equal(tkey[order], key)
    assert np.array_equal(val[order].view(np.uint32), val.view(np.uint32))
    sample = rng.choice(n, 48, replace=False)
    acc = oracle.vec_acc(pts.T, norm, rows=sample)
    for k, p in enumerate(sample):
        want = np.nonzero(acc[k] <= np.float32(thr))[0]
        assert np.array_equal(col[rp[p]:rp[p + 1]], want), p
        assert np.array_equal(val[rp[p]:rp[p + 1]].view(np.uint32), acc[k][want].view(np.uint32)), p


@pytest.mark.parametrize("name,meta", VEC_CLI_CASES, ids=[c[0] for c in VEC_CLI_CASES])
def test_vector_cli_matches_reference(name, meta, tmp_path):
    d = os.path.join(GOLDEN, "vectors", name)
    exe = os.path.join(PKG_DIR, "bin", meta["program"])
    assert os.path.exists(exe), f"{exe} not built"
    args = []
    for a in meta["argv"]:
        if a.startswith("IN:"):
            dst = tmp_path / ("in_" + a[3:])
            shutil.copyfile(golden_arg(a), dst)
            args.append(str(dst))
        else:
            args.append(golden_arg(a))
    print(f"vector cli case {name}: {meta['program']}", flush=True)
    r = subprocess.run([exe] + args, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == meta["returncode"], r.stderr.decode()[-2000:]
    assert r.stdout == open(os.path.join(d, "stdout.txt"), "rb").read()
    made = sorted(f for f in os.listdir(tmp_path) if not f.startswith("in_"))
    assert made == sorted(meta["files"])
    for fn in meta["files"]:
        assert (tmp_path / fn).read_bytes() == open(os.path.join(d, fn), "rb").read(), fn
