"""Serving loop (pra_server_*): requests admitted while others run overlap on
the GPU, and each result is bit-exact whether the request ran alone or in the
mix (and against the oracle on a subset)."""
import numpy as np
import pytest

import oracle as O
from programs import config4_instances, random_graph, tc_program

pytestmark = pytest.mark.gpu


def test_served_results_alone_and_in_mix():
    import pra_amd
    inst = config4_instances(count=16, seed=51, nmin=150, nmax=2500)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    print(f"[serve] {len(graphs)} requests", flush=True)
    alone = pra_amd.Server(nbuckets=4, slots=1, keep_rows=True)
    ref = []
    for e in graphs:
        t = alone.submit(e)
        r = alone.wait(t)
        assert r["status"] == 0
        ref.append((alone.rows(t), r["tuples"], r["joined"]))
    alone.close()
    # all at once into 4 slots under a budget that forces admission waits
    mix = pra_amd.Server(nbuckets=4, slots=4, hbm_budget=64 << 20, keep_rows=True)
    tickets = [mix.submit(e, est_bytes=int(200 * n * n)) for e, (n, _, _) in zip(graphs, inst)]
    res = [mix.wait(t) for t in tickets]
    for t, r, (rows, tuples, joined) in zip(tickets, res, ref):
        assert r["status"] == 0
        assert r["tuples"] == tuples and r["joined"] == joined
        assert np.array_equal(mix.rows(t), rows)
        assert r["t_submit"] <= r["t_start"] <= r["t_done"]
    # some requests ran concurrently
    spans = sorted((r["t_start"], r["t_done"]) for r in res)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    mix.close()
    for i in (0, 7):
        o = O.Oracle(4)
        To, *_ = tc_program(o, graphs[i])
        o.execute(1)
        assert np.array_equal(o.rows(To), ref[i][0])
