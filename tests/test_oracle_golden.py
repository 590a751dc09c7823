"""Pins the CPU oracle (oracle/pra_oracle.c) before it is trusted as the checker.

Known answers come from the reference itself (README.md:21 and the reference
run recorded with the ported data files), from textbook FNV-1a-64 vectors,
from the reference's own header-only tuple_hash compiled into oracle/_ref
where the reference tree is present, and from closed forms / brute force.
"""
import numpy as np
import pytest

import oracle as O
from programs import (DELTA, FULL, bfs_distance_histogram, small_random_graph, golden, grid_deltas, grid_graph,
                      grid_tc_size, layered_dag, nonlinear_tc_program, random_graph, sg_bruteforce, sg_program, tc_join_pairs, tc_pairs_bruteforce,
                      tc_program, two_rule_expected, two_rule_program, TWO_RULE_ORDERS, with_rowid, random_program)

TAG_T = 257

# FNV-1a-64 over LE bytes: textbook vectors (offset basis, prime 0x100000001b3)
FNV_KAT = [
    ([], 14695981039346656037),
    ([0], 12161962213042174405),
    ([1], 9929646806074584996),
    ([2], 16626593026977353223),
    ([3], 14394277620009763814),
    ([4], 3232700585171816769),
    ([2**64 - 1], 10157053723145373757),
    ([1, 2], 8581494755304202342),
    ([2, 1], 513837244993915590),
    ([123456789, 987654321, 42], 1489275174345521690),
]


def fnv_bytes(data):
    """FNV-1a-64 of a byte string (textbook definition)"""
    h = 0xcbf29ce484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv_known_answers():
    for words, want in FNV_KAT:
        assert O.tuple_hash(words) == want
        assert fnv_bytes(np.asarray(words, dtype="<u8").tobytes()) == want


def test_fnv_matches_reference_header():
    L = O.ref_tuple_hash_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes
    rng = np.random.default_rng(1)
    for n in range(0, 6):
        for _ in range(50):
            a = np.ascontiguousarray(rng.integers(0, 2**64 - 1, size=max(n, 1), dtype=np.uint64, endpoint=True))
            got = L.fnv_ref_tuple_hash(a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n)
            assert got == O.tuple_hash(a[:n])


def run_tc(nb, edges, batch_size=1, threshold=None):
    o = O.Oracle(nb)
    if threshold is not None:
        o.set_join_threshold(threshold)
    T, G, sfix, scopy = tc_program(o, edges)
    o.execute(batch_size)
    return o, T, G, sfix, scopy


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5])
def test_g5955_chain(nb):
    # README.md:23 claims 28; the reference run gives 4+3+2+1 = 10
    o, T, G, sfix, scopy = run_tc(nb, golden("g5955_edge_2_1_2"))
    assert o.count(T) == 10
    assert o.history(sfix) == [3, 2, 1, 0]
    assert o.history(scopy) == [4]


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5])
def test_g13236_reference_answers(nb):
    o, T, G, sfix, _ = run_tc(nb, golden("g13236_edge_2_1_2"))
    assert o.count(T) == 573                       # README.md:21
    assert o.count(G) == 60
    assert o.history(sfix) == [99, 132, 126, 82, 35, 19, 11, 7, 2, 0]
    assert o.joined() == 1140


def test_g13236_per_rank_split_k3():
    # per-rank DELTA row counts printed by the reference at 3 ranks
    o, T, G, sfix, _ = run_tc(3, golden("g13236_edge_2_1_2"))
    bh = o.bucket_history(sfix)[:, 1, :]          # SCC relation 1 is T
    want = {0: [37, 49, 44, 29, 17, 11, 9, 7, 2], 1: [26, 36, 31, 17, 8, 2], 2: [36, 47, 51, 36, 10, 6, 2]}
    for r, seq in want.items():
        got = [int(x) for x in bh[:, r] if x]
        assert got == seq


def test_g13236_arity_columns_rank_invariant():
    base = None
    for nb in [1, 2, 3, 5, 8]:
        o, T, *_ = run_tc(nb, golden("g13236_edge_2_1_2"))
        s = sorted(map(tuple, o.rows(T)[:, :2].tolist()))
        if base is None:
            base = s
        assert s == base
    pairs = [tuple(r) for r in golden("g13236_edge_2_1_2")[:, :2].tolist()]
    assert set(base) == {(z, x) for (x, z) in tc_pairs_bruteforce(pairs)}


def test_g5955_intern_ids_hand_derived():
    # k = 1: the copy seeds T(dst,src) in G's trie order, seq 0..3; then each
    # iteration interns in (delta row, G row) lexicographic emission order.
    o, T, *_ = run_tc(1, golden("g5955_edge_2_1_2"))
    base = TAG_T << 46
    want = {(1, 0): 0, (2, 1): 1, (3, 2): 2, (4, 3): 3, (2, 0): 4, (3, 1): 5, (4, 2): 6, (3, 0): 7, (4, 1): 8,
            (4, 0): 9}
    got = {(int(a), int(b)): int(i) for a, b, i in o.rows(T).tolist()}
    assert got == {k: base | v for k, v in want.items()}


def test_intern_id_fields():
    # tag in bits 46.., receiving bucket in bits 28..45, per-bucket seq dense from 0
    for nb in [1, 3, 5]:
        o, T, *_ = run_tc(nb, golden("g13236_edge_2_1_2"))
        rows = o.rows(T)
        ids = rows[:, 2].astype(object)
        assert all(int(i) >> 46 == TAG_T for i in ids)
        for r, i in zip(rows.tolist(), ids):
            assert (int(i) >> 28) & ((1 << 18) - 1) == O.tuple_hash([r[0]]) % nb
        for b in range(nb):
            seqs = sorted(int(i) & ((1 << 28) - 1) for i in ids if (int(i) >> 28) & ((1 << 18) - 1) == b)
            assert seqs == list(range(len(seqs)))


def test_intern_seq_follows_iteration_order():
    # tuples interned in an earlier iteration carry smaller seq numbers
    edges = golden("g13236_edge_2_1_2")
    o = O.Oracle(1)
    T, G, sfix, _ = tc_program(o, edges)
    o.execute(1)
    pairs = [tuple(r) for r in edges[:, :2].tolist()]
    total, deltas = bfs_distance_histogram(pairs)
    dist = {}
    for x, z in tc_pairs_bruteforce(pairs):
        pass
    rows = o.rows(T)
    seq = {(int(a), int(b)): int(i) & ((1 << 28) - 1) for a, b, i in rows.tolist()}
    # recompute shortest distances
    adj = {}
    for u, v in pairs:
        adj.setdefault(u, []).append(v)
    from collections import deque
    for x in adj:
        d = {}
        q = deque()
        for v in adj[x]:
            if v not in d:
                d[v] = 1
                q.append(v)
        while q:
            u = q.popleft()
            for v in adj.get(u, ()):
                if v not in d:
                    d[v] = d[u] + 1
                    q.append(v)
        for z, k in d.items():
            dist[(z, x)] = k
    order = sorted(seq, key=lambda t: seq[t])
    ks = [dist[t] for t in order]
    assert ks == sorted(ks)


@pytest.mark.parametrize("seed", range(8))
def test_small_random_bfs_histogram(seed):
    edges = small_random_graph(seed)
    pairs = [tuple(r) for r in edges[:, :2].tolist()]
    total, deltas = bfs_distance_histogram(pairs)
    for nb in [1, 3, 5]:
        o, T, G, sfix, _ = run_tc(nb, edges)
        assert o.count(T) == total
        assert o.history(sfix) == deltas
        assert o.joined() == tc_join_pairs(pairs)


def test_grid_closed_form():
    for n in [2, 5, 12, 30]:
        edges = grid_graph(n, seed=n)
        o, T, G, sfix, _ = run_tc(4, edges)
        assert o.count(T) == grid_tc_size(n)
        assert o.history(sfix) == grid_deltas(n)


def test_grid30_reference_prefix():
    # first fixed-point deltas of grid30 printed by the reference
    o, T, G, sfix, _ = run_tc(1, grid_graph(30))
    assert o.history(sfix)[:6] == [2521, 3244, 3910, 4520, 5075, 5576]
    assert o.count(T) == 215325


def test_batch_size_and_threshold_do_not_change_sets():
    edges = small_random_graph(3)
    o1, T1, *_ = run_tc(3, edges, batch_size=1)
    o2, T2, *_ = run_tc(3, edges, batch_size=4)
    o3, T3, *_ = run_tc(3, edges, batch_size=1, threshold=5)
    a = sorted(map(tuple, o1.rows(T1)[:, :2].tolist()))
    assert a == sorted(map(tuple, o2.rows(T2)[:, :2].tolist()))
    assert a == sorted(map(tuple, o3.rows(T3)[:, :2].tolist()))
    # batching only delays the test; ids are unchanged
    assert np.array_equal(o1.rows(T1), o2.rows(T2))


def test_threshold_rounds_reorder_interning():
    # with a tiny local-join threshold the join is split into rounds, so
    # round-1 arrivals from every source are interned before round-2 ones
    edges = golden("g13236_edge_2_1_2")
    o1, T1, *_ = run_tc(3, edges)
    o2, T2, *_ = run_tc(3, edges, threshold=3)
    assert sorted(map(tuple, o1.rows(T1)[:, :2].tolist())) == sorted(map(tuple, o2.rows(T2)[:, :2].tolist()))
    assert not np.array_equal(o1.rows(T1), o2.rows(T2))


@pytest.mark.parametrize("threshold", [None, 0, 1])
def test_threshold_rounds_hand_derived(threshold):
    # interning order of a two-rule program with and without split join
    # calls, derived by hand from the reference's round loop
    o = O.Oracle(1)
    if threshold is not None:
        o.set_join_threshold(threshold)
    out = two_rule_program(o)
    o.execute(1)
    assert np.array_equal(o.rows(out), two_rule_expected(TWO_RULE_ORDERS[threshold]))


def test_same_generation_bruteforce():
    for seed in range(3):
        edges = layered_dag(seed, nvert=120, nlayers=5, nedges=160)
        pairs = [tuple(r) for r in edges[:, :2].tolist()]
        want = sg_bruteforce(pairs)
        for nb in [1, 3]:
            o = O.Oracle(nb)
            h = sg_program(o, edges)
            o.execute(1)
            got = set(map(tuple, o.rows(h["SG"])[:, :2].tolist()))
            assert got == want


def test_nonlinear_tc_reaches_the_closure():
    # RIGHT joins (FULL x DELTA), joins against a changing full inner and an
    # acopy-maintained index relation still compute the transitive closure
    for seed in range(3):
        edges = random_graph(50, 90, seed)
        want = tc_pairs_bruteforce([tuple(r) for r in edges[:, :2].tolist()])
        for nb in [1, 3]:
            o = O.Oracle(nb)
            T, Tr, s = nonlinear_tc_program(o, edges)
            o.execute(1)
            got = set(map(tuple, o.rows(T)[:, :2].tolist()))
            assert got == want
            assert set(map(tuple, o.rows(Tr)[:, :2].tolist())) == {(y, x) for x, y in want}


def test_acopy_keeps_ids_and_dedups_on_all_columns():
    o = O.Oracle(2)
    A = o.rel_create(2, 1, 10)
    B = o.rel_create(2, 1, 11)
    rows = np.array([[1, 2, 100], [1, 2, 101], [3, 4, 102]], dtype=np.uint64)
    o.load_rows(A, rows, DELTA)
    s = o.scc_create(False, 1)
    o.scc_add_relation(s, A, False)
    o.scc_add_relation(s, B, False)
    o.add_acopy(s, B, A, DELTA, [1, 0, 2])
    o.execute(1)
    got = sorted(map(tuple, o.rows(B).tolist()))
    assert got == [(2, 1, 100), (2, 1, 101), (4, 3, 102)]


def test_copy_filter_neq():
    o = O.Oracle(1)
    A = o.rel_create(2, 1, 10)
    B = o.rel_create(2, 1, 11)
    o.load_rows(A, with_rowid([(1, 1), (1, 2), (3, 3), (4, 5)]), FULL)
    s = o.scc_create(False, 1)
    o.scc_add_relation(s, A, False)
    o.scc_add_relation(s, B, False)
    o.add_copy_filter(s, B, A, FULL, [1, 0], O.F_NEQ)
    o.execute(1)
    assert sorted(map(tuple, o.rows(B)[:, :2].tolist())) == [(2, 1), (5, 4)]


def test_empty_inputs():
    o, T, G, sfix, scopy = run_tc(3, np.zeros((0, 3), dtype=np.uint64))
    assert o.count(T) == 0
    assert o.history(sfix) == [0]
    assert o.history(scopy) == [0]


def test_random_programs_bucket_and_batch_invariance():
    # the same relation contents (data columns, as sets) for any bucket count
    # and batch size; at one bucket the ids do not depend on the batch size
    for seed in range(40):
        ref = None
        for nb, batch in [(1, 1), (3, 1), (5, 3)]:
            o = O.Oracle(nb)
            h = random_program(o, seed)
            o.execute(batch)
            sets = [sorted(map(tuple, o.rows(r)[:, :-1].tolist())) for r in h["rels"]]
            if ref is None:
                ref = sets
                ids1 = [o.rows(r) for r in h["rels"]]
            assert sets == ref, (seed, nb, batch)
        o = O.Oracle(1)
        h = random_program(o, seed)
        o.execute(4)
        assert all(np.array_equal(o.rows(r), x) for r, x in zip(h["rels"], ids1)), seed


def test_tc_size_condensation_matches_bruteforce():
    """the SCC-condensation |TC| used for checks of graphs of up to 50 K vertices equals brute
    force on small graphs with cycles, self-loops and isolated parts"""
    from programs import tc_size_condensation
    graphs = [small_random_graph(s)[:, :2] for s in range(4)]
    graphs += [random_graph(n, int(1.3 * n), 31 + n)[:, :2] for n in (50, 400, 1500)]
    graphs.append(np.array([[1, 1], [1, 2], [2, 3], [3, 2], [7, 8]], dtype=np.uint64))
    for p in graphs:
        assert tc_size_condensation(p)[0] == len(tc_pairs_bruteforce(p.tolist()))
