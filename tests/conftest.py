"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import importlib.util
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PKG_DIR = os.path.join(ROOT, "mnhn-tree-tools_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, PKG_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "mnhn_tree_tools_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mnhn_tree_tools_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import dbscan_oracle
    dbscan_oracle._lib()  # fail loudly if liboracle_sw.so is not built
    return dbscan_oracle


@pytest.fixture(scope="session")
def golden_pairs():
    with open(os.path.join(GOLDEN, "pairs_golden.json")) as f:
        return json.load(f)["pairs"]


@pytest.fixture(scope="session")
def golden_bruteforce():
    with open(os.path.join(GOLDEN, "bruteforce_ac6.json")) as f:
        return json.load(f)


def cli_cases():
    d = os.path.join(GOLDEN, "cli")
    out = []
    for name in sorted(os.listdir(d)):
        p = os.path.join(d, name, "cmd.json")
        if os.path.exists(p):
            with open(p) as f:
                out.append((name, json.load(f)))
    return out


def read_case_files(name, meta):
    d = os.path.join(GOLDEN, "cli", name)
    with open(os.path.join(d, "stdout.txt"), "rb") as f:
        stdout = f.read()
    files = {}
    for fn in meta["files"]:
        with open(os.path.join(d, fn), "rb") as f:
            files[fn] = f.read()
    return stdout, files


def vector_cases(programs=None):
    """(name, cmd.json) of the L1 / L2 vector-tool fixtures (and fasta2kmer)."""
    d = os.path.join(GOLDEN, "vectors")
    out = []
    for name in sorted(os.listdir(d)):
        with open(os.path.join(d, name, "cmd.json")) as f:
            meta = json.load(f)
        if programs is None or meta["program"] in programs:
            out.append((name, meta))
    return out


def golden_arg(a):
    """A fixture argument as the program received it: IN:<f> -> the input
    path, EPS:<f> -> the epsilon text stored in that input."""
    if a.startswith("IN:"):
        return os.path.join(GOLDEN, "inputs", a[3:])
    if a.startswith("EPS:"):
        with open(os.path.join(GOLDEN, "inputs", a[4:])) as f:
            return f.read().strip()
    return a
