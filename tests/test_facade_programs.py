"""Whole Slog-generated programs through the facade on the CPU oracle.

The reference's tests/worstcase-110-terms-4-m.cpp (60 relations, 19 SCCs, 39
joins, 15 copies, 31 acopies, 2 copy_filters; inputs = the reference's
data/worstcase-110-terms-4-m files, copied under tests/golden) is compiled
unchanged against include/parallel_RA_inc.h and linked with
oracle/libpra_abi_oracle.so.  Its printed total, 6,476,108 facts, is the
figure the reference itself prints for this program at 1 and 3 ranks."""
import json
import os
import sys

import pytest

from programs import GOLDEN, dump_digest, run_facade_program, total_facts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = "/root/reference/tests/worstcase-110-terms-4-m.cpp"
EXPECT = os.path.join(GOLDEN, "worstcase-110-terms-4-m.k3.json")
REFERENCE_TOTAL = 6476108


def test_worstcase_fixture_is_consistent():
    e = json.load(open(EXPECT))
    assert e["total_facts"] == REFERENCE_TOTAL
    assert len(e["relations"]) == 60
    assert sum(v[0] for v in e["relations"].values()) == REFERENCE_TOTAL
    # an input relation keeps the rows of its data file (app_2_57.dat: 330 rows)
    assert e["relations"]["rel_app_2_1_2"][0] == 330


@pytest.mark.skipif(not os.path.exists(SRC), reason="reference tree absent")
def test_worstcase_program_on_oracle(tmp_path):
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    exe = ge.build_reference_program(SRC, str(tmp_path / "wc_oracle"), libdir=ge.ORACLE, lib="pra_abi_oracle")
    r = run_facade_program(exe, "worstcase-110-terms-4-m", str(tmp_path / "run"), 3, str(tmp_path / "dump"))
    assert r.returncode == 0, r.stderr
    assert total_facts(r.stdout) == REFERENCE_TOTAL
    assert dump_digest(str(tmp_path / "dump")) == json.load(open(EXPECT))["relations"]


REF_TC = "/root/reference/tests/transitive_closure.cpp"
CKPT = os.path.join(GOLDEN, "tc_g13236_k2_checkpoints.json")


def _size_record(path):
    with open(path + ".size") as f:
        return [int(x) for x in f.read().split()]


@pytest.mark.skipif(not os.path.exists(REF_TC), reason="reference tree absent")
def test_tc_checkpoint_restart_on_oracle(tmp_path):
    # The reference at 2 ranks with a checkpoint every loop writes
    # checkpoint-1-5 holding |T full| = 534 and |T delta| = 19 and, restarted
    # from it, reaches 573 (measured on the reference itself).
    from programs import checkpoint_digest, dump_digest, run_tc_checkpointed
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    exe = ge.build_reference_program(REF_TC, str(tmp_path / "tc_oracle"), libdir=ge.ORACLE, lib="pra_abi_oracle")
    out, dump = run_tc_checkpointed(exe, str(tmp_path))
    ck = os.path.join(out, "checkpoint-1-5")
    full = [_size_record(os.path.join(ck, f"T_full_{b}")) for b in (0, 1)]
    delta = [_size_record(os.path.join(ck, f"T_delta_{b}")) for b in (0, 1)]
    assert sum(r for r, _ in full) == 534 and sum(r for r, _ in delta) == 19
    assert all(c == 3 for _, c in full + delta)
    assert open(os.path.join(ck, "scc_metadata")).read() == "2\n"
    # the copy SCC's own checkpoint and one per fixed-point loop (10 batches)
    assert sorted(os.listdir(out)) == ["checkpoint-1-%d" % i for i in range(10)] + ["checkpoint-2-0"]
    e = json.load(open(CKPT))
    assert checkpoint_digest(out) == e["checkpoints"]
    assert dump_digest(dump) == e["restart_dump"]
