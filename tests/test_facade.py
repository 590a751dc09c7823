"""The C++ facade (include/parallel_RA_inc.h) compiles the reference's own
programs unchanged and maps copy_filter lambdas to backend predicates."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TC = "/root/reference/tests/transitive_closure.cpp"


@pytest.mark.skipif(not os.path.exists(REF_TC), reason="reference tree absent")
def test_reference_tc_compiles_unchanged(tmp_path):
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    out = ge.build_reference_program(REF_TC, str(tmp_path / "TC"))
    assert os.path.exists(out)


def test_filter_lambda_classification(tmp_path):
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "facade_filter_probe.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    # EQ, NEQ, LT, GE, then predicates that are not a column comparison (one
    # that agrees with NEQ except at column 0 = 0, one that reads column 2),
    # then GT and LE
    assert out == ["0", "1", "2", "5", "-1", "-1", "-1", "4", "3"]


LOGGING_CASES = [("g5955", "g5955_edge_2_1_2"), ("tiny8", None)]


def logging_run(tmp_path, libdir, lib, edges_name):
    """the TC program with RAM::enable_logging on its fixed-point SCC, built
    against the facade and linked with `lib`; returns its relation dumps"""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_ref_vectors as M
    from programs import GOLDEN, write_bin
    exe = tmp_path / "logprobe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "facade_logging_probe.cpp"), "-L" + libdir, "-l" + lib,
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    if edges_name:
        edges = os.path.join(GOLDEN, edges_name)
    else:
        edges = str(tmp_path / "tiny8")
        write_bin(edges, M.tiny8_edges())
    env = dict(os.environ, PRA_BUCKETS="1")
    r = subprocess.run([str(exe), edges, os.path.join(GOLDEN, "empty_path_2_1_2")], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    return M.relation_dump_lines(r.stdout), r.stdout


@pytest.mark.parametrize("case,edges", LOGGING_CASES)
def test_enable_logging_prints_what_the_reference_prints(tmp_path, case, edges):
    """RAM::enable_logging: after every batch the facade prints each relation
    of the SCC exactly as the reference's relation::print does (golden lines
    from the reference itself at one rank, tests/golden/ref_log_tc_*_k1.txt);
    run here on the oracle-backed C ABI"""
    got, out = logging_run(tmp_path, os.path.join(ROOT, "oracle"), "pra_abi_oracle", edges)
    want = open(os.path.join(ROOT, "tests", "golden", f"ref_log_tc_{case}_k1.txt")).read().split("\n")[:-1]
    assert got == want
    # and the reference's per-iteration timing table header
    assert "Iteration#\tBuffer_creation_time\tComputation_time" in out
