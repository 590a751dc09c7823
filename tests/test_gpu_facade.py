"""Programs written against the reference's C++ surface run on the GPU through
the facade: the reference's own tests/transitive_closure.cpp (compiled
unchanged by build()) and the same-generation example program."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from programs import GOLDEN, layered_dag, sg_program, write_bin

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "parallel-ra_amd", "build")


def sizes(stdout):
    return {m.group(1): int(m.group(3)) for m in re.finditer(r"^(\S+): \{(\d+)\}\. \((\d+) total facts\)", stdout, re.M)}


@pytest.mark.skipif(not os.path.exists(os.path.join(BUILD, "TC_reference")), reason="reference TC not built")
@pytest.mark.parametrize("nb", [1, 3])
def test_reference_tc_program_on_gpu(tmp_path, nb):
    # 2-argument form: T is seeded from ../data/g5955/path_2_1_2 relative to cwd
    d = tmp_path / "data" / "g5955"
    d.mkdir(parents=True)
    shutil.copyfile(os.path.join(GOLDEN, "empty_path_2_1_2"), d / "path_2_1_2")
    shutil.copyfile(os.path.join(GOLDEN, "empty_path_2_1_2.size"), d / "path_2_1_2.size")
    cwd = tmp_path / "tests"
    cwd.mkdir()
    env = dict(os.environ, PRA_BUCKETS=str(nb))
    print(f"[facade] TC_reference nb={nb}", flush=True)
    r = subprocess.run([os.path.join(BUILD, "TC_reference"), os.path.join(GOLDEN, "g13236_edge_2_1_2")],
                       cwd=cwd, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = sizes(r.stdout)
    assert s["T"] == 573 and s["G"] == 60, r.stdout
    assert "Total facts across all relations 633" in r.stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(BUILD, "same_generation")), reason="example not built")
def test_same_generation_program_on_gpu(tmp_path):
    edges = layered_dag(9, nvert=2000, nlayers=8, nedges=2800)
    path = str(tmp_path / "edge_2_1_2")
    write_bin(path, edges)
    o = O.Oracle(2)
    h = sg_program(o, edges)
    o.execute(1)
    print("[facade] same_generation nb=2", flush=True)
    r = subprocess.run([os.path.join(BUILD, "same_generation"), path], env=dict(os.environ, PRA_BUCKETS="2"),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = sizes(r.stdout)
    assert s["sg_2_1_2"] == o.count(h["SG"])
    assert s["edge_2_1_2"] == 2800


@pytest.mark.skipif(not os.path.exists(os.path.join(BUILD, "worstcase-110-terms-4-m")),
                    reason="reference worstcase program not built")
def test_worstcase_program_on_gpu(tmp_path):
    # every relation's full version, ids included, bit-exact against the
    # oracle's run of the same unchanged program at 3 buckets
    import json
    from programs import dump_digest, run_facade_program, total_facts
    expect = json.load(open(os.path.join(GOLDEN, "worstcase-110-terms-4-m.k3.json")))
    print("[facade] worstcase-110-terms-4-m nb=3", flush=True)
    r = run_facade_program(os.path.join(BUILD, "worstcase-110-terms-4-m"), "worstcase-110-terms-4-m",
                           str(tmp_path / "run"), 3, str(tmp_path / "dump"), timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    print("\n".join(r.stdout.splitlines()[-8:]), flush=True)
    assert total_facts(r.stdout) == 6476108
    got = dump_digest(str(tmp_path / "dump"))
    bad = sorted(k for k in expect["relations"] if got.get(k) != expect["relations"][k])
    assert not bad, [(k, got.get(k), expect["relations"][k]) for k in bad[:5]]


@pytest.mark.skipif(not os.path.exists(os.path.join(BUILD, "TC_reference")), reason="reference TC not built")
def test_tc_checkpoint_restart_on_gpu(tmp_path):
    # every checkpoint file (per-rank full/delta rows with ids, size records,
    # scc_metadata) byte-identical to the oracle's run, and the restart from
    # checkpoint-1-5 reaching the same relations
    import json
    from programs import checkpoint_digest, dump_digest, run_tc_checkpointed
    e = json.load(open(os.path.join(GOLDEN, "tc_g13236_k2_checkpoints.json")))
    print("[facade] TC_reference checkpoint + restart nb=2", flush=True)
    out, dump = run_tc_checkpointed(os.path.join(BUILD, "TC_reference"), str(tmp_path))
    got = checkpoint_digest(out)
    assert sorted(got) == sorted(e["checkpoints"])
    for d in e["checkpoints"]:
        assert got[d] == e["checkpoints"][d], d
    assert dump_digest(dump) == e["restart_dump"]


@pytest.mark.parametrize("case,edges", [("g5955", "g5955_edge_2_1_2"), ("tiny8", None)])
def test_gpu_enable_logging_prints_what_the_reference_prints(tmp_path, case, edges):
    """RAM::enable_logging through the HIP library: the relation dumps after
    every batch equal the reference's own (tests/golden/ref_log_tc_*_k1.txt)"""
    from test_facade import logging_run
    got, _ = logging_run(tmp_path, os.path.join(ROOT, "parallel-ra_amd"), "pra_amd", edges)
    want = open(os.path.join(ROOT, "tests", "golden", f"ref_log_tc_{case}_k1.txt")).read().split("\n")[:-1]
    assert got == want
