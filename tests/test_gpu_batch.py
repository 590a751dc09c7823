"""Batched execution on random graphs (uniform random edges, vertex
counts log-uniform, 1.3 edges per vertex, uint64 ids): all instances as
program instances of one context run in lock step, every phase one batched
launch over all of them.  Per instance the rows (intern ids included), the
per-batch deltas and the join pairs equal the plain one-instance-per-context
run, |TC| equals the SCC-condensation count computed on the CPU, and a subset
equals the oracle.  The older host-thread form (pra_execute_many) is checked
against the plain run too."""
import numpy as np
import pytest

import oracle as O
from programs import log_uniform_instances, random_graph, tc_program, tc_size_condensation

pytestmark = pytest.mark.gpu


def plain_runs(graphs, nb):
    import pra_amd
    out = []
    for e in graphs:
        c = pra_amd.Context(nbuckets=nb, batch_size=1)
        T, G, sfix, _ = tc_program(c, e)
        c.execute()
        out.append((c.checksum(T), c.count(T), c.history(sfix), c.joined()))
        c.close()
    return out


@pytest.mark.parametrize("nb", [1, 8])
def test_lockstep_batch_equals_plain_and_closed_form(nb):
    import pra_amd
    inst = log_uniform_instances(count=24, seed=41, nmin=1000, nmax=8000)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    print(f"[batch] nb={nb} {len(graphs)} instances, sizes {[n for n, _, _ in inst]}", flush=True)
    plain = plain_runs(graphs, nb)
    c = pra_amd.Context(nbuckets=nb, batch_size=1)
    try:
        hs = []
        for i, e in enumerate(graphs):
            c.set_group(i)
            hs.append(tc_program(c, e))
        c.execute()
        for i, ((ck, cnt, hist, j), (T, G, sfix, _)) in enumerate(zip(plain, hs)):
            assert c.count(T) == cnt and c.checksum(T) == ck, i
            assert c.history(sfix) == hist, i
            assert c.group_joined(i) == j, i
            assert cnt == tc_size_condensation(graphs[i][:, :2])[0], i
        for i in (0, 13):
            o = O.Oracle(nb)
            To, Go, fo, _ = tc_program(o, graphs[i])
            o.execute(1)
            np.testing.assert_array_equal(c.rows(hs[i][0]), o.rows(To))
    finally:
        c.close()


def test_host_threads_equal_plain():
    import pra_amd
    inst = log_uniform_instances(count=12, seed=42, nmin=200, nmax=3000)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    plain = plain_runs(graphs, 8)
    ctxs, hs = [], []
    for e in graphs:
        c = pra_amd.Context(nbuckets=8, batch_size=1)
        hs.append(tc_program(c, e))
        ctxs.append(c)
    rc = pra_amd.execute_many(ctxs, concurrency=4)
    assert rc == [0] * len(ctxs)
    for (ck, cnt, hist, j), c, (T, G, sfix, _) in zip(plain, ctxs, hs):
        assert c.checksum(T) == ck and c.history(sfix) == hist and c.joined() == j
    for c in ctxs:
        c.close()
