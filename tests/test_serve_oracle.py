"""The serving loop (parallel-ra_amd/csrc/pra_server.cpp) linked with the CPU
oracle's C ABI (oracle/libpra_server_oracle.so): admission, preemption past
the HBM budget and recovery from a device that runs out of memory, checked on
the CPU.  Device bytes are modelled by the oracle ABI as the bytes of the rows
the instances hold, and PRA_ORACLE_DEVICE_BYTES makes a step fail with
PRA_E_NOMEM past a given size.  Every served result must equal the request's
result served alone and its SCC-condensation |TC|, however often the request
was restarted."""
import os

import numpy as np
import pytest

from programs import log_uniform_instances, random_graph, tc_size_condensation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVER_LIB = os.path.join(ROOT, "oracle", "libpra_server_oracle.so")
PRA_E_NOMEM = -8


@pytest.fixture(scope="module")
def stream():
    inst = log_uniform_instances(count=40, seed=77, nmin=30, nmax=400)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    return inst, graphs


@pytest.fixture(scope="module")
def library():
    import pra_amd
    if not os.path.exists(SERVER_LIB):
        pytest.skip("oracle/libpra_server_oracle.so not built (make -C oracle)")
    return pra_amd.load_library(SERVER_LIB)


@pytest.fixture(scope="module")
def alone(library, stream):
    import pra_amd
    _, graphs = stream
    os.environ.pop("PRA_ORACLE_DEVICE_BYTES", None)
    srv = pra_amd.Server(nbuckets=4, slots=1, keep_rows=True, library=library)
    out = []
    for e in graphs:
        t = srv.submit(e)
        r = srv.wait(t)
        assert r["status"] == 0 and r["restarts"] == 0
        out.append((r["tuples"], r["joined"], r["iterations"], srv.rows(t)))
    srv.close()
    assert [t for t, *_ in out] == [tc_size_condensation(e[:, :2])[0] for e in graphs]
    return out


def serve(library, graphs, est=0, device_bytes=None, **kw):
    import pra_amd
    if device_bytes is None:
        os.environ.pop("PRA_ORACLE_DEVICE_BYTES", None)
    else:
        os.environ["PRA_ORACLE_DEVICE_BYTES"] = str(int(device_bytes))
    try:
        srv = pra_amd.Server(nbuckets=4, keep_rows=True, library=library, **kw)
        tickets = [srv.submit(e, est_bytes=est) for e in graphs]
        res = [srv.wait(t) for t in tickets]
        rows = [srv.rows(t) if r["status"] == 0 else None for t, r in zip(tickets, res)]
        srv.close()
    finally:
        os.environ.pop("PRA_ORACLE_DEVICE_BYTES", None)
    return res, rows


def check_equal(res, rows, alone):
    for r, rw, (tuples, joined, iters, arows) in zip(res, rows, alone):
        assert r["status"] == 0
        assert (r["tuples"], r["joined"], r["iterations"]) == (tuples, joined, iters)
        assert np.array_equal(rw, arows)
        assert r["t_submit"] <= r["t_start"] <= r["t_done"]


def test_in_flight_together_equals_alone(library, stream, alone):
    _, graphs = stream
    res, rows = serve(library, graphs, slots=8)
    check_equal(res, rows, alone)
    assert sum(r["restarts"] for r in res) == 0
    spans = sorted((r["t_start"], r["t_done"]) for r in res)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_preemption_past_the_budget(library, stream, alone):
    """estimates of one byte admit up to 8 requests at once; a budget of half
    the largest result makes the loop release its newest admissions and serve
    them again later (one request always stays in flight)"""
    _, graphs = stream
    biggest = max(t for t, *_ in alone) * 24
    res, rows = serve(library, graphs, est=1, slots=8, hbm_budget=biggest // 2)
    check_equal(res, rows, alone)
    assert sum(r["restarts"] for r in res) > 0


def test_recovery_from_device_out_of_memory(library, stream, alone):
    """a modelled device that holds any one request but not eight: steps fail
    with PRA_E_NOMEM, the requests in flight are queued again on a fresh
    context with fewer in flight, and every request completes (alone, a TC
    request holds at most its result and its edges after any step: full and
    delta together never exceed the result)"""
    _, graphs = stream
    biggest = max(t for t, *_ in alone) * 24
    edges = max(len(e) for e in graphs) * 24
    res, rows = serve(library, graphs, slots=8, device_bytes=1.1 * (biggest + edges))
    check_equal(res, rows, alone)
    assert sum(r["restarts"] for r in res) > 0


def test_request_that_fails_alone_fails(library, stream, alone):
    """a request too big for the modelled device fails with PRA_E_NOMEM once
    it is alone; the others are served"""
    _, graphs = stream
    sizes = [t for t, *_ in alone]
    big = int(np.argmax(sizes))
    cap = 1.1 * (sorted(sizes)[-2] * 24 + max(len(e) for e in graphs) * 24)
    assert sizes[big] * 24 > cap
    res, rows = serve(library, graphs, slots=4, device_bytes=cap)
    for i, (r, rw, a) in enumerate(zip(res, rows, alone)):
        if i == big:
            assert r["status"] == PRA_E_NOMEM
        else:
            check_equal([r], [rw], [a])


def max_in_flight(res):
    """most requests between admission and completion at one time (an end
    at the same instant as a start comes first)"""
    ev = sorted([(r["t_start"], 1) for r in res] + [(r["t_done"], -1) for r in res], key=lambda e: (e[0], e[1]))
    cur = top = 0
    for _, d in ev:
        cur += d
        top = max(top, cur)
    return top


def test_admission_counts_requests_admitted_together(library, stream, alone):
    """requests admitted in the same pass count against the budget together:
    with every estimate a third of the budget, at most three are in flight at
    once although eight slots are free, and none is restarted (the real bytes
    stay far below the estimates)"""
    _, graphs = stream
    budget = 3 * 10**9
    res, rows = serve(library, graphs, est=budget // 3, slots=8, hbm_budget=budget)
    check_equal(res, rows, alone)
    assert sum(r["restarts"] for r in res) == 0
    assert max_in_flight(res) == 3


def test_large_requests_run_one_at_a_time(library, stream, alone):
    """requests whose estimate reaches PRA_SERVE_LARGE_BYTES (here: the
    server's own estimate for 200 vertices) are in flight one at a time;
    smaller ones arriving after a waiting large request are admitted before
    it, and every result equals the request served alone"""
    inst, graphs = stream
    # the server's own estimate from the edge count (1.3 edges per vertex)
    est = [int(0.18 * (len(e) / 1.3) ** 2 * 90) + (64 << 20) for e in graphs]
    bound = int(0.18 * 200 ** 2 * 90) + (64 << 20)
    large = [x >= bound for x in est]
    assert 2 <= sum(large) < len(graphs) - 2
    os.environ["PRA_SERVE_LARGE_BYTES"] = str(bound)
    try:
        res, rows = serve(library, graphs, slots=8)
    finally:
        os.environ.pop("PRA_SERVE_LARGE_BYTES", None)
    check_equal(res, rows, alone)
    spans = sorted((r["t_start"], r["t_done"]) for r, lg in zip(res, large) if lg)
    assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))
    overtaken = [(i, j) for i in range(len(graphs)) for j in range(i + 1, len(graphs))
                 if large[i] and not large[j] and res[j]["t_start"] < res[i]["t_start"]]
    assert overtaken
