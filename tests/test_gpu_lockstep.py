"""Several program instances in one context, run in lock step (one merged
iteration per step, every phase one batched launch over all instances): each
instance's rows (intern ids included), per-batch deltas and join pairs equal
the oracle's run of that instance alone."""
import numpy as np
import pytest

import oracle as O
from programs import (small_random_graph, random_graph, random_program, sg_program, layered_dag, tc_program,
                      two_rule_program)

pytestmark = pytest.mark.gpu


def oracle_run(k, build, bs=1):
    o = O.Oracle(k)
    h = build(o)
    o.execute(bs)
    return o, h


@pytest.mark.parametrize("k", [1, 3, 8])
def test_lockstep_tc_instances_equal_oracle(k):
    import pra_amd
    graphs = [random_graph(n, int(1.3 * n), 900 + n) for n in (50, 300, 120, 800, 60, 2000, 10)]
    graphs.append(small_random_graph(4))
    g = pra_amd.Context(nbuckets=k)
    try:
        handles = []
        for i, e in enumerate(graphs):
            g.set_group(i)
            handles.append(tc_program(g, e))
        g.execute()
        for i, (e, (T, G, sfix, scopy)) in enumerate(zip(graphs, handles)):
            print("instance", i, "k", k, flush=True)
            o, (To, Go, sfo, sco) = oracle_run(k, lambda b: tc_program(b, e))
            np.testing.assert_array_equal(g.rows(T), o.rows(To))
            assert g.history(sfix) == o.history(sfo)
            assert g.history(scopy) == o.history(sco)
            assert g.group_joined(i) == o.joined()
    finally:
        g.close()


@pytest.mark.parametrize("k", [1, 3])
def test_lockstep_mixed_programs_equal_oracle(k):
    """instances with different SCC structures: TC, same generation, a
    two-rule program and random multi-SCC programs"""
    import pra_amd
    builders = [
        lambda b: tc_program(b, small_random_graph(7)),
        lambda b: sg_program(b, layered_dag(5, nvert=300, nlayers=5, nedges=420)),
        lambda b: two_rule_program(b),
        lambda b: random_program(b, 11),
        lambda b: tc_program(b, random_graph(500, 650, 77)),
        lambda b: random_program(b, 12),
    ]

    class Rec:
        """records the handles a builder creates, in creation order"""

        def __init__(self, b):
            self.b, self.rels, self.sccs = b, [], []

        def __getattr__(self, n):
            return getattr(self.b, n)

        def rel_create(self, *a):
            h = self.b.rel_create(*a)
            self.rels.append(h)
            return h

        def scc_create(self, *a):
            h = self.b.scc_create(*a)
            self.sccs.append(h)
            return h

    g = pra_amd.Context(nbuckets=k)
    try:
        recs = []
        for i, bld in enumerate(builders):
            g.set_group(i)
            r = Rec(g)
            bld(r)
            recs.append(r)
        g.execute()
        for i, bld in enumerate(builders):
            print("instance", i, "k", k, flush=True)
            o = O.Oracle(k)
            ro = Rec(o)
            bld(ro)
            o.execute(1)
            for a, b in zip(recs[i].rels, ro.rels):
                np.testing.assert_array_equal(g.rows(a), o.rows(b))
            for a, b in zip(recs[i].sccs, ro.sccs):
                assert g.history(a) == o.history(b)
            assert g.group_joined(i) == o.joined()
    finally:
        g.close()


def test_lockstep_many_instances_equal_oracle():
    """100 instances (past the 64 rules and outputs the batched arrival plan
    takes as kernel arguments, so it plans from tables in device memory)"""
    import pra_amd
    graphs = [random_graph(20 + 7 * i, int(1.3 * (20 + 7 * i)), 4000 + i) for i in range(100)]
    g = pra_amd.Context(nbuckets=3)
    try:
        handles = []
        for i, e in enumerate(graphs):
            g.set_group(i)
            handles.append(tc_program(g, e))
        g.execute()
        for i in list(range(0, 100, 9)) + [99]:
            T, G, sfix, scopy = handles[i]
            o, (To, Go, sfo, sco) = oracle_run(3, lambda b: tc_program(b, graphs[i]))
            np.testing.assert_array_equal(g.rows(T), o.rows(To))
            assert g.history(sfix) == o.history(sfo)
            assert g.group_joined(i) == o.joined()
    finally:
        g.close()


def test_step_with_trim_and_abandoned_instance():
    """pra_ctx_step with the scratch given back after every step
    (pra_ctx_trim), an instance abandoned in the middle of its fixed point
    (pra_ctx_release_group of a running instance) and a new program built in
    its place: the other instances and the new one equal the oracle"""
    import pra_amd
    k = 3
    graphs = [random_graph(n, int(1.3 * n), 4100 + n) for n in (400, 1500, 90, 2500)]
    late = random_graph(700, 910, 4242)
    g = pra_amd.Context(nbuckets=k, batch_size=1)
    try:
        handles = {}
        for i, e in enumerate(graphs):
            g.set_group(i)
            handles[i] = (e, tc_program(g, e))
        steps = 0
        while True:
            active = g.step()
            g.trim()
            steps += 1
            if steps == 5:
                assert not g.group_status(1)[0]
                g.release_group(1)          # abandoned mid-run
                del handles[1]
                g.set_group(1)
                handles[1] = (late, tc_program(g, late))
            if not active and steps > 5:
                break
        for i, (e, (T, G, sfix, scopy)) in sorted(handles.items()):
            print("instance", i, flush=True)
            o, (To, Go, sfo, sco) = oracle_run(k, lambda b: tc_program(b, e))
            np.testing.assert_array_equal(g.rows(T), o.rows(To))
            assert g.history(sfix) == o.history(sfo)
            assert g.group_joined(i) == o.joined()
    finally:
        g.close()
