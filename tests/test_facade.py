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
    # EQ, NEQ, LT, GE, and a predicate that is not a column comparison
    assert out == ["0", "1", "2", "5", "-1"]
