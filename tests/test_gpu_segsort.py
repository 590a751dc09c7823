"""Newts too large for the top-byte bins of the packed-key sort: the
segmented sort (two counting passes over the keys' top bytes, then one LDS
sort per segment) and the byte-pass sort it falls back to.  A lowered bin
capacity (PRA_SORT_TEST_BIN_CAP) sends small programs down both paths, which
must equal the oracle, ids and per-batch deltas included; at full size the
largest instance of the 256-graph batch (49,691 vertices) gives the same rows down both paths."""
import os

import numpy as np
import pytest

import oracle as O
from programs import (log_uniform_instances, layered_dag, random_graph, random_program, sg_program, tc_program,
                      tc_size_condensation)

pytestmark = pytest.mark.gpu


def context(k, env):
    import pra_amd
    old = {key: os.environ.get(key) for key in env}
    os.environ.update(env)
    try:
        return pra_amd.Context(nbuckets=k, batch_size=1)
    finally:
        for key, v in old.items():
            if v is None:
                os.environ.pop(key, None)
            else:
                os.environ[key] = v


class Rec:
    def __init__(self, b):
        self.b, self.rels, self.sccs = b, [], []

    def __getattr__(self, n):
        return getattr(self.b, n)

    def rel_create(self, *a):
        r = self.b.rel_create(*a)
        self.rels.append(r)
        return r

    def scc_create(self, *a):
        s = self.b.scc_create(*a)
        self.sccs.append(s)
        return s


BUILDERS = [
    ("tc_random800", lambda b: tc_program(b, random_graph(800, 1040, 801))),
    ("sg_dag", lambda b: sg_program(b, layered_dag(7, nvert=400, nlayers=5, nedges=560))),
    ("random_program_3", lambda b: random_program(b, 3)),
    ("random_program_8", lambda b: random_program(b, 8)),
]


@pytest.mark.parametrize("path", ["segmented", "byte_passes"])
@pytest.mark.parametrize("k", [1, 3])
@pytest.mark.parametrize("name,build", BUILDERS)
def test_large_newt_paths_equal_oracle(name, build, k, path):
    env = {"PRA_SORT_TEST_BIN_CAP": "16", "PRA_SEG_SORT": "1" if path == "segmented" else "0"}
    print(f"[segsort] {name} k={k} {path}", flush=True)
    g = context(k, env)
    try:
        rg = Rec(g)
        build(rg)
        g.execute()
        o = O.Oracle(k)
        ro = Rec(o)
        build(ro)
        o.execute(1)
        for a, b in zip(rg.rels, ro.rels):
            np.testing.assert_array_equal(g.rows(a), o.rows(b))
        for a, b in zip(rg.sccs, ro.sccs):
            assert g.history(a) == o.history(b)
        assert g.joined() == o.joined()
    finally:
        g.close()


def test_largest_batch_instance_both_paths():
    import time
    n, m, sd = max(log_uniform_instances(count=256), key=lambda i: i[0])
    e = random_graph(n, m, sd)
    out = {}
    for path in ("segmented", "byte_passes"):
        print(f"[segsort] n={n} {path}", flush=True)
        g = context(8, {"PRA_SEG_SORT": "1" if path == "segmented" else "0"})
        try:
            T, G, sfix, _ = tc_program(g, e)
            g.synchronize()
            t0 = time.perf_counter()
            g.execute()
            g.synchronize()
            dt = time.perf_counter() - t0
            out[path] = (g.count(T), g.checksum(T), g.history(sfix), g.joined())
            print(f"[segsort] n={n} {path}: {dt:.3f} s", flush=True)
        finally:
            g.close()
    assert out["segmented"] == out["byte_passes"]
    assert out["segmented"][0] == tc_size_condensation(e[:, :2])[0]
