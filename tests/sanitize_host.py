"""Host DBSCAN and split-set code under AddressSanitizer + UBSan (CPU only).

Builds dbscan.c and cluster_io.c with gcc -fsanitize=address,undefined into a
temporary library that resolves the engine symbols from libmnhn_sw.so, then
runs the CPU parity tests of those files against it in a child process with
the sanitizer runtimes preloaded.  Not collected by pytest; run it directly:

    python tests/sanitize_host.py

Exit status 0 when every test passes with no sanitizer report.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "mnhn-tree-tools_amd")

TESTS = ["test_dbscan_over_csr_matches_oracle", "test_dbscan_over_vcsr_matches_oracle",
         "test_dbscan_dense_neighbourhoods_match_oracle", "test_dbscan_sweep_matches_oracle",
         "test_dbscan_vcsr_sweep_matches_oracle", "test_dbscan_sweep_long_lists_outside_a_cluster",
         "test_dbscan_passes_over_whole_lists_match_oracle", "test_dbscan_threaded_pass_lists_match_oracle",
         "test_dbscan_parallel_row_scans_match_oracle"]

CHILD = r"""
import ctypes, inspect, os, sys
sys.path[:0] = [{here!r}, {root!r}, os.path.join({root!r}, "oracle")]
import conftest, dbscan_oracle as oracle, test_abi_cpu as t
pkg = conftest.load_pkg()
lib = ctypes.CDLL({lib!r})
pkg.load_library = lambda *a, **k: lib
class Env:
    def setenv(self, k, v):
        os.environ[k] = v
for name in {tests!r}:
    f = getattr(t, name)
    f(*[{{"pkg": pkg, "oracle": oracle, "monkeypatch": Env()}}[p] for p in inspect.signature(f).parameters])
    for k in ("MNHN_SWEEP_INDEX", "MNHN_HOST_THREADS"):
        os.environ.pop(k, None)
    print("ok", name, flush=True)
lib.generate_split_set_relation.restype = ctypes.POINTER(t.Conn)
lib.generate_split_set_relation.argtypes = [t.SplitSet, t.SplitSet]
t.test_relation_of_disjoint_split_sets((lib, None), oracle)
t.test_relation_and_sort_side_effect((lib, None), oracle)
print("ok relations", flush=True)
"""


def main():
    with tempfile.TemporaryDirectory() as tmp:
        lib = os.path.join(tmp, "libmnhn_host_asan.so")
        src = [os.path.join(PKG, "csrc", "host", f) for f in ("dbscan.c", "cluster_io.c")]
        subprocess.run(["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=undefined", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include")]
                       + src + ["-o", lib, "-L" + PKG, "-lmnhn_sw", "-Wl,-rpath," + PKG, "-lpthread", "-lm"],
                       check=True)
        rt = [subprocess.run(["gcc", "-print-file-name=" + r], capture_output=True, text=True,
                             check=True).stdout.strip() for r in ("libasan.so", "libubsan.so")]
        env = dict(os.environ, LD_PRELOAD=":".join(rt), ASAN_OPTIONS="detect_leaks=0")
        code = CHILD.format(here=HERE, root=ROOT, lib=lib, tests=TESTS)
        return subprocess.run([sys.executable, "-c", code], env=env).returncode


if __name__ == "__main__":
    sys.exit(main())
