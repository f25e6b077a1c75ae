"""GPU integration tests: the drop-in CLIs produce byte-identical stdout and
output files to the compiled reference on every golden case."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, PKG_DIR, cli_cases, read_case_files

pytestmark = pytest.mark.gpu

CASES = cli_cases()


def _run(name, meta, tmp_path, env_extra=None):
    stdout, files = read_case_files(name, meta)
    exe = os.path.join(PKG_DIR, "bin", meta["program"])
    assert os.path.exists(exe), f"{exe} not built"
    fasta = tmp_path / "in.fasta"
    shutil.copyfile(os.path.join(GOLDEN, "inputs", meta["input"]), fasta)
    if meta["program"] == "cluster_dbscan_SW":
        argv = [exe, str(fasta)] + meta["args"] + [meta["prefix"]]
    else:
        argv = [exe, str(fasta)] + [a if a != "PREFIX" else meta["prefix"] for a in meta["args"]]
    env = dict(os.environ)
    env.update(env_extra or {})
    print(f"cli case {name}: {' '.join(argv[:1])}", flush=True)
    r = subprocess.run(argv, cwd=tmp_path, capture_output=True, env=env, timeout=300)
    assert r.returncode == meta["returncode"], r.stderr.decode()[-2000:]
    assert r.stdout == stdout
    got = sorted(f for f in os.listdir(tmp_path) if f.startswith(meta["prefix"]))
    assert got == sorted(files)
    for fn, data in files.items():
        assert (tmp_path / fn).read_bytes() == data, fn
    return r


@pytest.mark.parametrize("name,meta", CASES, ids=[c[0] for c in CASES])
def test_cli_matches_reference(name, meta, tmp_path):
    _run(name, meta, tmp_path)


@pytest.mark.parametrize("name", ["dbscan_test2", "dbscan_c1small_e400", "adaptive_testxs33"])
def test_cli_per_query_launch_mode(name, tmp_path):
    """The reference-shaped flow: one one-vs-all launch per region query."""
    meta = dict(cli_cases())[name]
    _run(name, meta, tmp_path, {"MNHN_REGION_QUERY": "launch"})


@pytest.mark.parametrize("name,cap", [("adaptive_c1small", "0"), ("adaptive_c1small", "1000"),
                                      ("adaptive_c1small", "20000"), ("dbscan_c1full_e400", "0")])
def test_cli_lists_beyond_memory_fall_back(name, cap, tmp_path):
    """Epsilon lists that do not fit (forced with MNHN_SW_LIST_MAX_ENTRIES)
    switch the run to one one-vs-all launch per region query instead of
    exiting, as the reference recomputes a row per query: stdout and files
    stay byte-identical.  On the 320-sequence sweep (eps 100 + 60 k) a cap of
    1000 keeps the lists at eps 100 and falls back from eps 160 on, 20000
    keeps the lists up to eps 580 and falls back at eps 640."""
    meta = dict(cli_cases())[name]
    r = _run(name, meta, tmp_path, {"MNHN_SW_LIST_MAX_ENTRIES": cap})
    assert b"region queries recompute their distances" in r.stderr


def test_split_set_no_truncate(tmp_path):
    """A pre-existing longer split-set file keeps its tail bytes, as the
    reference's open(O_WRONLY | O_CREAT) leaves it."""
    name = "adaptive_c1small"
    meta = dict(cli_cases())[name]
    stdout, files = read_case_files(name, meta)
    first = sorted(files)[0]
    junk = b"X" * (len(files[first]) + 37)
    (tmp_path / first).write_bytes(junk)
    exe = os.path.join(PKG_DIR, "bin", meta["program"])
    fasta = tmp_path / "in.fasta"
    shutil.copyfile(os.path.join(GOLDEN, "inputs", meta["input"]), fasta)
    argv = [exe, str(fasta)] + [a if a != "PREFIX" else meta["prefix"] for a in meta["args"]]
    r = subprocess.run(argv, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0
    data = (tmp_path / first).read_bytes()
    assert data[:len(files[first])] == files[first]
    assert data[len(files[first]):] == junk[len(files[first]):]


@pytest.mark.parametrize("name", ["dbscan_c1small_e400", "adaptive_testxs33"])
def test_cli_persisted_eps_lists(name, tmp_path):
    """MNHN_ECSR: the first run computes and writes the lists, the second reads
    them instead of running Smith-Waterman; both outputs stay byte-identical to
    the reference's; a file from other data is not used."""
    meta = dict(cli_cases())[name]
    path = str(tmp_path.parent / f"{name}.ecsr")
    for k, msg in ((0, b"written to"), (1, b"read from")):
        run_dir = tmp_path / f"run{k}"
        run_dir.mkdir()
        stdout, files = read_case_files(name, meta)
        exe = os.path.join(PKG_DIR, "bin", meta["program"])
        fasta = run_dir / "in.fasta"
        shutil.copyfile(os.path.join(GOLDEN, "inputs", meta["input"]), fasta)
        if meta["program"] == "cluster_dbscan_SW":
            argv = [exe, str(fasta)] + meta["args"] + [meta["prefix"]]
        else:
            argv = [exe, str(fasta)] + [a if a != "PREFIX" else meta["prefix"] for a in meta["args"]]
        env = dict(os.environ, MNHN_ECSR=path)
        print(f"ecsr case {name} run {k}", flush=True)
        r = subprocess.run(argv, cwd=run_dir, capture_output=True, env=env, timeout=300)
        assert r.returncode == meta["returncode"]
        assert msg in r.stderr, r.stderr.decode()[-500:]
        assert r.stdout == stdout
        for fn, data in files.items():
            assert (run_dir / fn).read_bytes() == data, fn
    # other data (one sequence dropped): the file does not match, lists are recomputed
    other = tmp_path / "other.fasta"
    seqs = open(os.path.join(GOLDEN, "inputs", meta["input"]), "rb").read().split(b">")[1:]
    other.write_bytes(b"".join(b">" + s for s in seqs[:-1]))
    exe = os.path.join(PKG_DIR, "bin", "cluster_dbscan_SW")
    r = subprocess.run([exe, str(other), "400", "5", "o"], cwd=tmp_path, capture_output=True,
                       env=dict(os.environ, MNHN_ECSR=path), timeout=300)
    assert r.returncode == 0 and b"written to" in r.stderr


# ---- downstream: the reference's split_sets_to_newick on our outputs --------------

NEWICK_DIR = os.path.join(GOLDEN, "newick")
NEWICK_CASES = sorted(os.listdir(NEWICK_DIR)) if os.path.isdir(NEWICK_DIR) else []
NEWICK_TOOL = os.path.join(os.path.dirname(PKG_DIR), "oracle", "_ref", "split_sets_to_newick")


@pytest.mark.parametrize("name", NEWICK_CASES)
def test_newick_tool_reads_our_outputs(name, tmp_path):
    """The split sets and stdout of the drop-in adaptive CLIs feed the
    reference's split_sets_to_newick (compiled from its source) to the same
    tree as the reference's own outputs do."""
    if not os.path.exists(NEWICK_TOOL):
        pytest.skip("reference split_sets_to_newick not built (make -C oracle ref)")
    import json
    meta = json.load(open(os.path.join(NEWICK_DIR, name, "cmd.json")))
    src = json.load(open(os.path.join(GOLDEN, meta["source"], "cmd.json")))
    exe = os.path.join(PKG_DIR, "bin", src["program"])
    if "argv" in src:  # vector fixture: argv with IN:/EPS: inputs
        from conftest import golden_arg
        args = [golden_arg(a) for a in src["argv"]]
    else:
        fasta = tmp_path / "in.fasta"
        shutil.copyfile(os.path.join(GOLDEN, "inputs", src["input"]), fasta)
        args = [str(fasta)] + [a if a != "PREFIX" else src["prefix"] for a in src["args"]]
    print(f"newick case {name}: {src['program']}", flush=True)
    r = subprocess.run([exe] + args, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    (tmp_path / "stdout.txt").write_bytes(r.stdout)
    argv = [NEWICK_TOOL] + ["stdout.txt" if a == "STDOUT" else a for a in meta["args"]] + meta["sets"]
    t = subprocess.run(argv, cwd=tmp_path, capture_output=True, timeout=120)
    assert t.returncode == meta["returncode"]
    assert t.stdout == open(os.path.join(NEWICK_DIR, name, "stdout.txt"), "rb").read()


# ---- the scalar seam: the reference's own CLIs with score() bound to mnhn_score ----

REF_DIR_BUILT = os.path.join(os.path.dirname(PKG_DIR), "oracle", "_ref")
SCORE_CASES = [c for c in cli_cases() if c[0] in ("dbscan_test2", "dbscan_test3", "dbscan_test4",
                                                  "adaptive_test3")]


@pytest.mark.parametrize("name,meta", SCORE_CASES, ids=[c[0] for c in SCORE_CASES])
def test_reference_cli_with_score_bound_to_engine(name, meta, tmp_path):
    """The reference's cluster_dbscan_SW / adaptive_clustering_SW compiled from
    its sources with score() defined by csrc/binding/reference_score.c (one
    GPU pair per call through mnhn_score) reproduce the reference's outputs."""
    exe = os.path.join(REF_DIR_BUILT, meta["program"] + "_mnhn_score")
    if not os.path.exists(exe):
        pytest.skip("relinked reference CLIs not built (make -C oracle relink)")
    stdout, files = read_case_files(name, meta)
    fasta = tmp_path / "in.fasta"
    shutil.copyfile(os.path.join(GOLDEN, "inputs", meta["input"]), fasta)
    if meta["program"] == "cluster_dbscan_SW":
        argv = [exe, str(fasta)] + meta["args"] + [meta["prefix"]]
    else:
        argv = [exe, str(fasta)] + [a if a != "PREFIX" else meta["prefix"] for a in meta["args"]]
    print(f"relinked reference case {name}", flush=True)
    r = subprocess.run(argv, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == meta["returncode"], r.stderr.decode()[-2000:]
    assert r.stdout == stdout
    for fn, data in files.items():
        assert (tmp_path / fn).read_bytes() == data, fn


LIB_CASES = [c for c in cli_cases() if c[0] in ("dbscan_c1full_e400", "dbscan_c1small_e400", "dbscan_c1small_e160",
                                                "dbscan_test2")]


@pytest.mark.parametrize("name,meta", LIB_CASES, ids=[c[0] for c in LIB_CASES])
def test_reference_cli_linked_against_library_alone(name, meta, tmp_path):
    """The reference's cluster_dbscan_SW main (CPU build, calling
    dbscan_SW(dataset, float, int), dbscan.h:30) compiled from its source and
    linked with -lmnhn_sw only -- no reference object -- reproduces the
    reference's stdout and cluster files."""
    exe = os.path.join(REF_DIR_BUILT, "cluster_dbscan_SW_mnhn_lib")
    if not os.path.exists(exe):
        pytest.skip("library-linked reference CLI not built (make -C oracle relink)")
    stdout, files = read_case_files(name, meta)
    fasta = tmp_path / "in.fasta"
    shutil.copyfile(os.path.join(GOLDEN, "inputs", meta["input"]), fasta)
    print(f"library-linked reference case {name}", flush=True)
    r = subprocess.run([exe, str(fasta)] + meta["args"] + [meta["prefix"]], cwd=tmp_path, capture_output=True,
                       timeout=300)
    assert r.returncode == meta["returncode"], r.stderr.decode()[-2000:]
    assert r.stdout == stdout
    got = sorted(f for f in os.listdir(tmp_path) if f.startswith(meta["prefix"]))
    assert got == sorted(files)
    for fn, data in files.items():
        assert (tmp_path / fn).read_bytes() == data, fn


def test_mnhn_score_matches_golden_pairs(pkg, golden_pairs):
    """mnhn_score, the scalar seam, against the reference's score() vectors."""
    import ctypes
    lib = pkg.load_library()
    lib.mnhn_score.restype = ctypes.c_uint
    lib.mnhn_score.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t,
                               ctypes.c_void_p]
    for a, b, d in golden_pairs[:120]:
        a, b = a.encode(), b.encode()
        assert lib.mnhn_score(a, b, len(a), len(b), None) == d, (a[:20], b[:20])
