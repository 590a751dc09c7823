"""Serving loop (pra_server_*) with continuous batching: requests in flight
advance together one lock-step iteration at a time and queued requests join
between iterations as slots and memory free up.  Each result is bit-exact
whether the request ran alone or in the mix, equals the SCC-condensation |TC|,
and a subset equals the oracle."""
import time

import numpy as np
import pytest

import oracle as O
from programs import log_uniform_instances, random_graph, tc_program, tc_size_condensation

pytestmark = pytest.mark.gpu


def test_served_results_alone_and_in_mix():
    import pra_amd
    inst = log_uniform_instances(count=16, seed=51, nmin=150, nmax=2500)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    print(f"[serve] {len(graphs)} requests", flush=True)
    alone = pra_amd.Server(nbuckets=4, slots=1, keep_rows=True)
    ref = []
    for e in graphs:
        t = alone.submit(e)
        r = alone.wait(t)
        assert r["status"] == 0
        ref.append((alone.rows(t), r["tuples"], r["joined"], r["iterations"]))
    alone.close()
    # all at once into 4 slots under a budget that forces admission waits
    mix = pra_amd.Server(nbuckets=4, slots=4, hbm_budget=64 << 20, keep_rows=True)
    tickets = [mix.submit(e, est_bytes=int(200 * n * n)) for e, (n, _, _) in zip(graphs, inst)]
    res = [mix.wait(t) for t in tickets]
    for t, r, (rows, tuples, joined, iters) in zip(tickets, res, ref):
        assert r["status"] == 0
        assert r["tuples"] == tuples and r["joined"] == joined and r["iterations"] == iters
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


def test_poisson_stream_continuous_batching():
    """a seeded Poisson stream of 300 requests at 1,000 req/s (sizes from the
    log-uniform size distribution, reduced range) into 16 slots under an HBM budget:
    every result equals its SCC-condensation |TC| and its alone run's tuple
    count and join pairs; requests joined while others were in flight"""
    import pra_amd
    inst = log_uniform_instances(count=300, seed=55, nmin=200, nmax=4000)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    want = [tc_size_condensation(e[:, :2])[0] for e in graphs]
    rng = np.random.default_rng(1000)
    gaps = rng.exponential(1.0 / 1000.0, size=len(graphs))
    srv = pra_amd.Server(nbuckets=8, slots=16, hbm_budget=8 << 30)
    t0 = time.perf_counter()
    due = 0.0
    tickets = []
    for e, gap in zip(graphs, gaps):
        due += gap
        now = time.perf_counter() - t0
        if due > now:
            time.sleep(due - now)
        tickets.append(srv.submit(e))
    res = [srv.wait(t) for t in tickets]
    srv.close()
    print("[serve stream] p50/p99 ms", np.percentile([r["t_done"] - r["t_submit"] for r in res], [50, 99]) * 1e3,
          flush=True)
    assert all(r["status"] == 0 for r in res)
    assert [r["tuples"] for r in res] == want
    # some admitted while others were in flight
    starts = sorted(r["t_start"] for r in res)
    spans = sorted((r["t_start"], r["t_done"]) for r in res)
    assert sum(1 for i in range(len(spans) - 1) if spans[i + 1][0] < spans[i][1]) > len(res) // 4
    alone = pra_amd.Server(nbuckets=8, slots=1)
    for i in (0, 77, 150, 299):
        r = alone.wait(alone.submit(graphs[i]))
        assert (r["tuples"], r["joined"], r["iterations"]) == (res[i]["tuples"], res[i]["joined"], res[i]["iterations"])
    alone.close()
    del starts
