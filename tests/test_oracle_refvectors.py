"""Pins the oracle (oracle/pra_oracle.c) against vectors the reference itself
printed (tests/golden/ref_vectors.json, written by
tests/golden/make_ref_vectors.py from the reference built in oracle/_ref and
run under mpirun on k ranks).

Per program, rank count k and batch size the oracle must reproduce, bit for
bit: every batch's global delta (RAM::check_for_fixed_point), every batch's
per-bucket delta of every SCC relation (rank r = bucket r), and every
relation's final full version in the reference's row order, intern-id column
included.  The ids at k > 1 follow the reference's arrival order -- receiver
bucket by source bucket, then rule, then emission (src/RAM/RA_tasks.cpp:
831-835, src/comm/all_to_allv_comm.cpp:64-142) -- which these vectors fix."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_ref_vectors as MRV  # noqa: E402

with open(os.path.join(HERE, "golden", "ref_vectors.json")) as f:
    VECTORS = json.load(f)
BUILDERS = {name: build for name, build, _, _, _ in MRV.cases()}


def check_against_reference(backend, handles, want, bs, per_bucket=None):
    """compare a run backend (after execute) with one reference vector;
    handles = (relation handles by recorder index, scc handles by recorder index)"""
    rels, sccs = handles
    for s in want["sccs"]:
        got = backend.history(sccs[s["scc"]])
        assert got == s["history"], (s["scc"], got, s["history"])
        if per_bucket is not None and bs == 1:
            pb = per_bucket(sccs[s["scc"]])
            assert [[list(map(int, r)) for r in it] for it in pb] == s["delta_per_rank"], s["scc"]
    for r, f in want["final"].items():
        rows = np.ascontiguousarray(backend.rows(rels[int(r)]), dtype="<u8")
        assert rows.shape[0] == f["rows_count"], (r, rows.shape[0], f["rows_count"])
        if "rows" in f:
            np.testing.assert_array_equal(rows, np.array(f["rows"], dtype=np.uint64).reshape(rows.shape),
                                          err_msg=f"relation {r}")
        assert hashlib.sha256(rows.tobytes()).hexdigest() == f["sha256"], r


class Handles:
    """wraps a backend to record the handles in creation order"""

    def __init__(self, b):
        self.b = b
        self.rels = []
        self.sccs = []

    def __getattr__(self, name):
        return getattr(self.b, name)

    def rel_create(self, *a):
        h = self.b.rel_create(*a)
        self.rels.append(h)
        return h

    def scc_create(self, *a):
        h = self.b.scc_create(*a)
        self.sccs.append(h)
        return h


def run_case(backend, key):
    name, k, bs = key.split("/")
    hb = Handles(backend)
    BUILDERS[name](hb)
    return hb, int(bs[1:])


@pytest.mark.parametrize("key", sorted(VECTORS))
def test_oracle_matches_reference_vectors(key):
    k = int(key.split("/")[1][1:])
    o = O.Oracle(k)
    hb, bs = run_case(o, key)
    o.execute(bs)
    check_against_reference(o, (hb.rels, hb.sccs), VECTORS[key], bs, per_bucket=o.bucket_history)


def test_vectors_cover_cross_bucket_arrivals():
    """the k > 1 id vectors include receiving buckets fed by several source
    buckets and by both rules of one iteration (two_rule_wide), so an oracle
    that interned rule-major or by destination order would fail"""
    from programs import FULL  # noqa: F401
    A = [(y, 10 * (y % 4)) for y in range(1, 13)]
    B = [(y, 500 + y % 3) for y in range(1, 13)] + [(y, 600) for y in range(2, 13, 2)]
    C = [(y, 500 + (y + 1) % 3) for y in range(1, 13)] + [(y, 700) for y in range(1, 13, 3)]
    for k in (2, 3):
        feeds = {}
        for rule, inner in enumerate((B, C)):
            for y, x in A:
                src = O.tuple_hash([y]) % k
                for y2, z in inner:
                    if y2 == y:
                        feeds.setdefault(O.tuple_hash([z]) % k, set()).add((src, rule))
        assert any(len({s for s, _ in v}) >= 2 and len({r for _, r in v}) == 2 for v in feeds.values()), k
        assert f"two_rule_wide/k{k}/b1" in VECTORS
