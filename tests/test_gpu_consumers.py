"""GPU integration tests of the reference's other score() consumers
(find_closest_sequence_SW, find_satellite, compareSW, silhouette): stdout and
output files byte-identical to the compiled reference's on the golden cases,
plus the host-side matrix statistics against numpy."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR

pytestmark = pytest.mark.gpu

CASE_DIR = os.path.join(GOLDEN, "consumers")
CASES = sorted(d for d in os.listdir(CASE_DIR) if os.path.exists(os.path.join(CASE_DIR, d, "cmd.json")))


@pytest.mark.parametrize("name", CASES)
def test_consumer_matches_reference(name, tmp_path):
    d = os.path.join(CASE_DIR, name)
    meta = json.load(open(os.path.join(d, "cmd.json")))
    exe = os.path.join(PKG_DIR, "bin", meta["program"])
    assert os.path.exists(exe), f"{exe} not built"
    args = []
    for a in meta["argv"]:
        if a.startswith("IN:"):
            dst = tmp_path / ("in_" + a[3:])
            shutil.copyfile(os.path.join(GOLDEN, "inputs", a[3:]), dst)
            args.append(str(dst))
        else:
            args.append(a)
    print(f"consumer case {name}: {meta['program']}", flush=True)
    r = subprocess.run([exe] + args, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == meta["returncode"], r.stderr.decode()[-2000:]
    assert r.stdout == open(os.path.join(d, "stdout.txt"), "rb").read()
    for fn in meta["files"]:
        assert (tmp_path / fn).read_bytes() == open(os.path.join(d, fn), "rb").read(), fn
