"""Parity of the HIP backend (through the C ABI) with the CPU oracle.

Integer set arithmetic: every comparison is bit-exact, including the intern-id
column and the per-iteration delta counts.  Full-size inputs are checked
through size-independent properties (closed-form sizes and delta sequences,
sortedness, uniqueness, id-field structure).
"""
import numpy as np
import pytest

import oracle as O
from programs import (DELTA, FULL, bfs_distance_histogram, small_random_graph, golden, grid_deltas, grid_graph,
                      grid_tc_size, layered_dag, nonlinear_tc_program, random_graph, sg_bruteforce, sg_program, tc_program,
                      two_rule_expected, two_rule_program, TWO_RULE_ORDERS, with_rowid)

pytestmark = pytest.mark.gpu


def gpu_ctx(nb, batch_size=1):
    import pra_amd
    return pra_amd.Context(nbuckets=nb, batch_size=batch_size)


def both(nb, build, batch_size=1):
    o = O.Oracle(nb)
    g = gpu_ctx(nb, batch_size)
    ho = build(o)
    hg = build(g)
    o.execute(batch_size)
    g.execute()
    return o, ho, g, hg


def assert_same_rel(o, ro, g, rg, what=""):
    a = o.rows(ro)
    b = g.rows(rg)
    assert a.shape == b.shape, f"{what}: {a.shape} vs {b.shape}"
    assert np.array_equal(a, b), f"{what}: rows differ"


def check_tc(nb, edges, batch_size=1):
    print(f"[tc] nb={nb} edges={len(edges)} batch={batch_size}", flush=True)
    o, (To, Go, fo, co), g, (Tg, Gg, fg, cg) = both(nb, lambda b: tc_program(b, edges), batch_size)
    assert_same_rel(o, To, g, Tg, "T")
    assert_same_rel(o, Go, g, Gg, "G")
    assert o.history(fo) == g.history(fg)
    assert o.history(co) == g.history(cg)
    assert o.joined() == g.joined()
    return o, g, To, Tg, fg


@pytest.mark.parametrize("nb", [1, 2, 3, 5, 8])
def test_tc_golden_graphs(nb):
    o, g, To, Tg, fg = check_tc(nb, golden("g13236_edge_2_1_2"))
    assert g.count(Tg) == 573
    assert g.history(fg) == [99, 132, 126, 82, 35, 19, 11, 7, 2, 0]
    o, g, To, Tg, fg = check_tc(nb, golden("g5955_edge_2_1_2"))
    assert g.count(Tg) == 10


@pytest.mark.parametrize("seed", range(6))
def test_tc_small_random(seed):
    edges = small_random_graph(seed)
    pairs = [tuple(r) for r in edges[:, :2].tolist()]
    total, deltas = bfs_distance_histogram(pairs)
    for nb in [1, 3, 5]:
        o, g, To, Tg, fg = check_tc(nb, edges)
        assert g.count(Tg) == total
        assert g.history(fg) == deltas


def test_tc_batch_size():
    check_tc(3, small_random_graph(11), batch_size=3)


def test_tc_empty_and_tiny():
    check_tc(3, np.zeros((0, 3), dtype=np.uint64))
    check_tc(2, with_rowid([(5, 5)]))              # self loop
    check_tc(4, with_rowid([(1, 2), (2, 1)]))      # 2-cycle


@pytest.mark.parametrize("nb", [1, 8])
def test_tc_grid_random_ids(nb):
    n = 30
    o, g, To, Tg, fg = check_tc(nb, grid_graph(n, seed=2024))
    assert g.count(Tg) == grid_tc_size(n)
    assert g.history(fg) == grid_deltas(n)


def test_tc_random_graph_medium():
    check_tc(8, random_graph(1500, 1950, seed=5))


def test_same_generation_parity():
    edges = layered_dag(4, nvert=600, nlayers=6, nedges=840)
    pairs = [tuple(r) for r in edges[:, :2].tolist()]
    want = sg_bruteforce(pairs)
    for nb in [1, 3]:
        print(f"[sg] nb={nb}", flush=True)
        o, ho, g, hg = both(nb, lambda b: sg_program(b, edges))
        for k in ["E", "I", "SG", "A", "B"]:
            assert_same_rel(o, ho[k], g, hg[k], k)
        for k in ["base", "filter", "rec"]:
            assert o.history(ho[k]) == g.history(hg[k])
        assert set(map(tuple, g.rows(hg["SG"])[:, :2].tolist())) == want


def test_acopy_and_filter_parity():
    rows = np.array([[1, 2, 100], [1, 2, 101], [3, 4, 102], [7, 7, 103]], dtype=np.uint64)

    def build(b):
        A = b.rel_create(2, 1, 10)
        B = b.rel_create(2, 1, 11)
        C = b.rel_create(2, 1, 12)
        b.load_rows(A, rows, DELTA)
        s = b.scc_create(False, 1)
        b.scc_add_relation(s, A, False)
        b.scc_add_relation(s, B, False)
        b.scc_add_relation(s, C, False)
        b.add_acopy(s, B, A, DELTA, [1, 0, 2])
        b.add_copy_filter(s, C, A, DELTA, [1, 0], O.F_NEQ)
        return A, B, C

    for nb in [1, 2]:
        o, ho, g, hg = both(nb, build)
        for i in range(3):
            assert_same_rel(o, ho[i], g, hg[i], f"rel{i}")


def test_grid150_tc_properties():
    # the full-size config: 150x150 grid, uint64 ids, 8 virtual buckets
    n, nb = 150, 8
    print("[grid150] grid150 nb=8", flush=True)
    g = gpu_ctx(nb)
    T, G, sfix, scopy = tc_program(g, grid_graph(n, seed=150))
    g.execute()
    assert g.count(T) == grid_tc_size(n) == 128233125
    assert g.history(sfix) == grid_deltas(n)
    assert g.joined() == 253069050
    rows = g.rows(T)
    # bucket-major, lexicographic within bucket, no duplicate keys
    b = np.array([O.tuple_hash([int(x)]) % nb for x in rows[::100000, 0]])
    assert np.all(np.diff(b) >= 0)
    ids = rows[:, 2]
    assert np.all((ids >> np.uint64(46)) == np.uint64(257))
    seq = ids & np.uint64((1 << 28) - 1)
    bucket = (ids >> np.uint64(28)) & np.uint64((1 << 18) - 1)
    for bb in range(nb):
        s = np.sort(seq[bucket == bb])
        assert np.array_equal(s, np.arange(s.size, dtype=np.uint64))


def test_same_generation_dag10k_full_size():
    # 10,000-vertex 12-layer DAG, 14,000 uint64 edges, 8 virtual buckets:
    # 1,392,703 sg tuples; every relation's rows (ids included) and deltas vs the oracle
    edges = layered_dag(3)
    print("[sg_dag10k] same generation nb=8", flush=True)
    o, ho, g, hg = both(8, lambda b: sg_program(b, edges))
    for k in ["E", "I", "SG", "A", "B"]:
        assert_same_rel(o, ho[k], g, hg[k], k)
    for k in ["base", "filter", "rec"]:
        assert o.history(ho[k]) == g.history(hg[k])
    assert g.count(hg["SG"]) == 1392703
    assert o.joined() == g.joined()


def _mix64(x):
    M = (1 << 64) - 1
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & M
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & M
    x ^= x >> 33
    return x


def _keyhash(k):
    M = (1 << 64) - 1
    h = 0x9E3779B97F4A7C15 ^ len(k)
    for i, v in enumerate(k):
        h = _mix64(h ^ ((v + 0x632BE59BD9B4E019 * (i + 1)) & M))
    return h or 1


def test_rows_checksum_definition():
    g = gpu_ctx(3)
    T, G, sfix, _ = tc_program(g, golden("g13236_edge_2_1_2"))
    g.execute()
    rows = g.rows(T).tolist()
    M = (1 << 64) - 1
    want = 0
    for i, r in enumerate(rows):
        want = (want + _mix64(_keyhash([int(x) for x in r]) ^ _mix64((i + 0x2545F4914F6CDD1D) & M))) & M
    assert g.checksum(T) == want


def test_intern_ids_flowing_into_key_columns():
    # the projection keeps the delta row's intern id as a key column, so new
    # values appear every iteration: exercises the packed-sort-key fallback
    edges = layered_dag(6, nvert=300, nlayers=6, nedges=420)

    def build(b):
        T = b.rel_create(2, 1, 257)
        G = b.rel_create(2, 1, 256)
        b.load_rows(G, edges, FULL)
        fix = b.scc_create(True, 1)
        b.scc_add_relation(fix, G, False)
        b.scc_add_relation(fix, T, True)
        b.add_join(fix, T, T, DELTA, G, FULL, [3, 2])
        seed = b.scc_create(False, 2)
        b.scc_add_relation(seed, G, False)
        b.scc_add_relation(seed, T, True)
        b.add_copy(seed, T, G, FULL, [1, 2])
        b.add_dependency(seed, fix)
        return T, fix

    for nb in [1, 3]:
        o, (To, fo), g, (Tg, fg) = both(nb, build)
        assert_same_rel(o, To, g, Tg, "T")
        assert o.history(fo) == g.history(fg)


@pytest.mark.parametrize("nb", [1, 3])
def test_nonlinear_tc_parity(nb):
    # joins whose inner full version changes every iteration (binary-search
    # probes, merge-maintained full) and a FULL x DELTA (RIGHT) join
    for seed in range(2):
        edges = random_graph(400, 700, 40 + seed)
        print(f"[nonlinear tc] nb={nb} seed={seed}", flush=True)
        o, (To, Tro, so), g, (Tg, Trg, sg) = both(nb, lambda b: nonlinear_tc_program(b, edges))
        assert_same_rel(o, To, g, Tg, "T")
        assert_same_rel(o, Tro, g, Trg, "Tr")
        assert o.history(so) == g.history(sg)


@pytest.mark.parametrize("nb", [1, 3])
def test_threshold_rounds_parity(nb):
    # join calls over the local-join threshold split into rounds, as the
    # reference's local_join / execute_in_batches do: ids, per-batch deltas and
    # join pairs equal to the oracle's, which restates the rounds exactly
    cases = [("g13236 tc", golden("g13236_edge_2_1_2"), tc_program, [0, 3, 17]),
             ("random tc", random_graph(300, 520, 7), tc_program, [17, 200]),
             ("same generation", layered_dag(5, nvert=300, nlayers=6, nedges=420), sg_program, [9, 60])]
    for name, edges, prog, thresholds in cases:
        for t in thresholds:
            print(f"[threshold] {name} nb={nb} threshold={t}", flush=True)

            def build(b):
                b.set_join_threshold(t)
                h = prog(b, edges)
                return h if isinstance(h, dict) else {"T": h[0], "G": h[1], "fix": h[2], "seed": h[3]}
            o, ho, g, hg = both(nb, build)
            for k in ho:
                if isinstance(ho[k], int) and k not in ("fix", "seed", "base", "filter", "rec"):
                    assert_same_rel(o, ho[k], g, hg[k], f"{name} t={t} {k}")
            for k in ("fix", "seed", "base", "filter", "rec"):
                if k in ho:
                    assert o.history(ho[k]) == g.history(hg[k]), (name, t, k)
            assert o.joined() == g.joined()


@pytest.mark.parametrize("threshold", [None, 0, 1])
def test_threshold_rounds_hand_derived(threshold):
    # the interning order derived by hand from the reference's round loop
    print(f"[threshold hand] threshold={threshold}", flush=True)
    g = gpu_ctx(1)
    if threshold is not None:
        g.set_join_threshold(threshold)
    out = two_rule_program(g)
    g.execute()
    assert np.array_equal(g.rows(out), two_rule_expected(TWO_RULE_ORDERS[threshold]))


@pytest.mark.parametrize("nb", [1, 3])
def test_skewed_keys_take_the_byte_pass_sort(nb):
    # every tuple the copy and the join produce shares its first column (a
    # hub, then a sink), so all packed sort keys fall into one top-byte bin,
    # which overflows its capacity: the newt sort takes the one-sweep byte
    # passes instead of the bin sort, with the same result as the oracle
    print(f"[skewed bins] nb={nb}", flush=True)
    hub, sink = 1_000_000, 2_000_000
    edges = with_rowid([(x, hub) for x in range(3000)] + [(hub, sink)])
    g = gpu_ctx(nb)
    g.set_profiling(True)
    Tg, Gg, fg, cg = tc_program(g, edges)
    g.execute()
    o = O.Oracle(nb)
    To, Go, fo, co = tc_program(o, edges)
    o.execute(1)
    assert_same_rel(o, To, g, Tg, "T")
    assert o.history(fo) == g.history(fg)
    assert o.count(To) == 3000 + 1 + 3000
    ks = g.kernel_stats()
    assert ks["bin_sort"]["launches"] == 0 and ks["radix_scatter"]["launches"] > 0


@pytest.mark.parametrize("nb", [1, 3])
@pytest.mark.parametrize("n", [2, 3, 100, 200, 1500, 3000, 20000])
def test_wide_rows_sort_paths(nb, n):
    # five-column rows (one column of random 64-bit values, four of 24-bit
    # ones): the acopy's rows (six columns, the id moved) and the interned
    # copy's (once their dictionary ranks do not fit one packed 64-bit key)
    # are sorted column by column, by the one-block network (n <= 2048), the
    # one-block radix passes (n <= 16384) or the one-sweep byte passes; both
    # must order them as the oracle does
    print(f"[wide sort] nb={nb} n={n}", flush=True)
    rng = np.random.default_rng(n * 10 + nb)
    vals = rng.integers(0, 2**24, size=(n, 5), dtype=np.uint64)
    vals[:, 2] = rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    rows = np.concatenate([vals, np.arange(7000, 7000 + n, dtype=np.uint64)[:, None]], axis=1)

    def build(b):
        A = b.rel_create(5, 1, 300)
        B = b.rel_create(5, 2, 301)
        C = b.rel_create(5, 1, 302)
        b.load_rows(A, rows, FULL)
        s = b.scc_create(True, 40)
        b.scc_add_relation(s, A, True)
        b.scc_add_relation(s, B, False)
        b.scc_add_relation(s, C, False)
        b.add_acopy(s, B, A, DELTA, [3, 0, 4, 1, 2, 5])
        b.add_copy(s, C, A, DELTA, [2, 4, 0, 3, 1])
        return A, B, C, s

    g = gpu_ctx(nb)
    g.set_profiling(True)
    Ag, Bg, Cg, sg = build(g)
    g.execute()
    o = O.Oracle(nb)
    Ao, Bo, Co, so = build(o)
    o.execute(1)
    assert_same_rel(o, Bo, g, Bg, "acopy target")
    assert_same_rel(o, Co, g, Cg, "copy target")
    assert o.history(so) == g.history(sg)
    ks = g.kernel_stats()
    if n <= 1500:
        assert ks["small_sort"]["launches"] > 0
    elif n <= 16384:
        assert ks["block_sort"]["launches"] > 0
    else:
        assert ks["radix_scatter"]["launches"] > 0
    g.close()


@pytest.mark.parametrize("nb", [1, 3])
def test_scc_with_more_rules_than_one_launch_holds(nb):
    # 95 rules over 64 relations in one fixed-point SCC (the most an SCC may
    # hold): more rules than one batched count-and-scan launch (36) or one
    # arrival-plan launch (64) takes, so the first is split and the second
    # falls back to one plan per output; 31 acopy targets sort in one batch
    print(f"[many rules] nb={nb}", flush=True)
    rng = np.random.default_rng(70 + nb)
    pairs = sorted({(int(a), int(b)) for a, b in rng.integers(0, 40, size=(300, 2))})
    rows = np.array([[a, b, 9000 + k] for k, (a, b) in enumerate(pairs)], dtype=np.uint64)

    def build(b):
        A = b.rel_create(2, 1, 400)
        b.load_rows(A, rows, FULL)
        s = b.scc_create(True, 41)
        b.scc_add_relation(s, A, True)
        outs = []
        for i in range(31):
            B = b.rel_create(2, 1, 401 + i)
            b.scc_add_relation(s, B, False)
            b.add_acopy(s, B, A, DELTA, [1, 0, 2] if i % 2 else [0, 1, 2])
            outs.append(B)
        for i in range(32):
            C = b.rel_create(2, 1, 501 + i)
            b.scc_add_relation(s, C, False)
            b.add_copy(s, C, A, DELTA, [1, 0])
            b.add_copy(s, C, A, DELTA, [0, 1] if i % 3 else [1, 0])
            outs.append(C)
        return outs, s

    g = gpu_ctx(nb)
    g.set_profiling(True)
    og, sg = build(g)
    g.execute()
    o = O.Oracle(nb)
    oo, so = build(o)
    o.execute(1)
    for i, (ro, rg) in enumerate(zip(oo, og)):
        assert_same_rel(o, ro, g, rg, f"output {i}")
    assert o.history(so) == g.history(sg)
    assert g.kernel_stats()["small_sort"]["launches"] >= 1
    g.close()


@pytest.mark.parametrize("nb", [1, 3])
def test_candidate_overflow_rerun_in_a_batched_iteration(nb):
    # two outputs per iteration (the batched claim / decision / partition
    # path): A = reachability over 100 sources feeding a 5-node chain into Z,
    # B = A's new pairs keyed on Z joined with Z's 2,000 out-edges.  B's
    # candidates jump from 2,000 to 200,000 in the iteration where the pairs
    # (Z, source) appear, past their capacity (twice the last count, at least
    # 65,536): the output is re-run at its exact size, with the same result as
    # the oracle
    print(f"[overflow rerun] nb={nb}", flush=True)
    chain = [1000, 1001, 1002, 1003, 1004]        # c1 .. c4, Z
    g1 = [(x, chain[0]) for x in range(100)] + list(zip(chain[:-1], chain[1:]))
    g2 = [(chain[-1], 5000 + j) for j in range(2000)]

    def build(b):
        G1 = b.rel_create(2, 1, 700)
        G2 = b.rel_create(2, 1, 701)
        A = b.rel_create(2, 1, 702)
        B = b.rel_create(2, 1, 703)
        b.load_rows(G1, with_rowid(g1), FULL)
        b.load_rows(G2, with_rowid(g2), FULL)
        fix = b.scc_create(True, 50)
        for r, mv in ((G1, False), (G2, False), (A, True), (B, False)):
            b.scc_add_relation(fix, r, mv)
        b.add_join(fix, A, A, DELTA, G1, FULL, [3, 1])
        b.add_join(fix, B, A, DELTA, G2, FULL, [3, 1])
        seed = b.scc_create(False, 51)
        b.scc_add_relation(seed, G1, False)
        b.scc_add_relation(seed, A, True)
        b.add_copy(seed, A, G1, FULL, [1, 0])
        b.add_dependency(seed, fix)
        return A, B, fix

    g = gpu_ctx(nb)
    g.set_profiling(True)
    Ag, Bg, fg = build(g)
    g.execute()
    # one batched claim per iteration, and one more for the re-run
    assert g.kernel_stats()["dedup_claim"]["launches"] == len(g.history(fg)) + 1
    o = O.Oracle(nb)
    Ao, Bo, fo = build(o)
    o.execute(1)
    assert_same_rel(o, Ao, g, Ag, "A")
    assert_same_rel(o, Bo, g, Bg, "B")
    assert o.history(fo) == g.history(fg)
    assert o.count(Bo) > 100 * 2000
    g.close()
