"""Batched execution (pra_execute_many: independent fixed points overlapping
on one GPU) is bit-exact per instance against the plain one-instance-at-a-time
path, and both against the oracle on a subset."""
import numpy as np
import pytest

import oracle as O
from programs import config4_instances, random_graph, tc_program

pytestmark = pytest.mark.gpu


def test_batched_equals_plain_and_oracle():
    import pra_amd
    inst = config4_instances(count=12, seed=41, nmin=200, nmax=3000)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    print(f"[batch] {len(graphs)} instances, sizes {[n for n, _, _ in inst]}", flush=True)
    plain = []
    for e in graphs:
        c = pra_amd.Context(nbuckets=8, batch_size=1)
        T, G, sfix, _ = tc_program(c, e)
        c.execute()
        plain.append((c.rows(T), c.history(sfix), c.joined()))
        c.close()
    ctxs, hs = [], []
    for e in graphs:
        c = pra_amd.Context(nbuckets=8, batch_size=1)
        hs.append(tc_program(c, e))
        ctxs.append(c)
    rc = pra_amd.execute_many(ctxs, concurrency=4)
    assert rc == [0] * len(ctxs)
    for (rows, hist, j), c, (T, G, sfix, _) in zip(plain, ctxs, hs):
        assert np.array_equal(c.rows(T), rows)
        assert c.history(sfix) == hist
        assert c.joined() == j
    for i in (0, 5, 11):
        o = O.Oracle(8)
        To, Go, fo, _ = tc_program(o, graphs[i])
        o.execute(1)
        assert np.array_equal(o.rows(To), plain[i][0])
        assert o.history(fo) == plain[i][1]
    for c in ctxs:
        c.close()
