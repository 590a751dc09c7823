"""Programs, inputs and closed forms shared by the tests and the benchmark.

Programs are written against a backend object with the program-building
calls of the oracle (oracle/oracle.py) and of the HIP binding
(parallel-ra_amd/pra_amd.py), both mirroring the reference's C++ surface.
"""
import os
import random
from collections import deque

import numpy as np

DELTA, FULL = 0, 1
F_NEQ = 1
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------ tsv_to_bin ----

def read_bin(path):
    """tsv_to_bin format: <path> = rows of LE u64, <path>.size = "rows\\ncols"
    (utility/tsv_to_bin.cpp, src/IO/parallel_io.cpp:115-235)."""
    with open(path + ".size") as f:
        rows, cols = (int(x) for x in f.read().split())
    if rows == 0:
        return np.zeros((0, cols), dtype=np.uint64)
    a = np.fromfile(path, dtype="<u8")
    return a.reshape(rows, cols).astype(np.uint64)


def write_bin(path, rows):
    rows = np.ascontiguousarray(np.asarray(rows, dtype="<u8"))
    rows.tofile(path)
    with open(path + ".size", "w") as f:
        f.write(f"{rows.shape[0]}\n{rows.shape[1]}")


def with_rowid(pairs):
    """(src, dst) pairs -> tsv_to_bin rows (src, dst, rowid)."""
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    return np.column_stack([p, np.arange(p.shape[0], dtype=np.uint64)])


def golden(name):
    return read_bin(os.path.join(GOLDEN, name))


# -------------------------------------------------------------- programs ----

def tc_program(b, edges, seed_rows=None):
    """tests/transitive_closure.cpp:56-77 on backend `b`: returns (T, G, scc_fix, scc_copy)."""
    T = b.rel_create(2, 1, 257)
    G = b.rel_create(2, 1, 256)
    if seed_rows is not None and len(seed_rows):
        b.load_rows(T, seed_rows, FULL)
    b.load_rows(G, edges, FULL)
    s13237 = b.scc_create(True, 13237)
    b.scc_add_relation(s13237, G, False)
    b.scc_add_relation(s13237, T, True)
    b.add_join(s13237, T, T, DELTA, G, FULL, [3, 1])
    s13239 = b.scc_create(False, 13239)
    b.scc_add_relation(s13239, G, False)
    b.scc_add_relation(s13239, T, True)
    b.add_copy(s13239, T, G, FULL, [1, 0])
    b.add_dependency(s13239, s13237)
    return T, G, s13237, s13239


def sg_program(b, edges):
    """Same generation over `edges` (p, x, id):
         sg(x,y) :- edge(p,x), edge(p,y), x != y.
         sg(x,y) :- edge(a,x), sg(a,b), edge(b,y).
    The recursive rule is a join chain re-keyed by an acopy:
         A(x,b)    <- sg(a,b) JOIN edge(a,x)          on a
         B(b,x,id) <- acopy of A                       (re-key on b)
         sg(x,y)   <- B(b,x) JOIN edge(b,y)            on b
    Returns dict of relation handles and SCC handles."""
    E = b.rel_create(2, 1, 300)
    I = b.rel_create(2, 1, 301)     # edge(p,x) JOIN edge(p,y) -> (x, y)
    SG = b.rel_create(2, 1, 302)
    A = b.rel_create(2, 1, 303)
    B = b.rel_create(2, 1, 304)
    b.load_rows(E, edges, FULL)
    s_base = b.scc_create(False, 1)
    b.scc_add_relation(s_base, E, False)
    b.scc_add_relation(s_base, I, False)
    b.add_join(s_base, I, E, FULL, E, FULL, [1, 3])
    s_filter = b.scc_create(False, 2)
    b.scc_add_relation(s_filter, I, True)
    b.scc_add_relation(s_filter, SG, False)
    b.add_copy_filter(s_filter, SG, I, DELTA, [0, 1], F_NEQ)
    s_rec = b.scc_create(True, 3)
    b.scc_add_relation(s_rec, E, False)
    b.scc_add_relation(s_rec, SG, True)
    b.scc_add_relation(s_rec, A, False)
    b.scc_add_relation(s_rec, B, False)
    b.add_join(s_rec, A, SG, DELTA, E, FULL, [3, 1])
    b.add_acopy(s_rec, B, A, DELTA, [1, 0, 2])
    b.add_join(s_rec, SG, B, DELTA, E, FULL, [1, 3])
    b.add_dependency(s_base, s_filter)
    b.add_dependency(s_filter, s_rec)
    return {"E": E, "I": I, "SG": SG, "A": A, "B": B, "base": s_base, "filter": s_filter, "rec": s_rec}


# ------------------------------------------------------------ generators ----

def small_random_graph(seed, n=40, m=60):
    """40 vertices, 60 distinct edges uniform without self-loops, random.Random(seed)."""
    rng = random.Random(seed)
    edges = []
    seen = set()
    while len(edges) < m:
        u, v = rng.randrange(n), rng.randrange(n)
        if u != v and (u, v) not in seen:
            seen.add((u, v))
            edges.append((u, v))
    return with_rowid(edges)


def distinct_u64(n, seed):
    rng = np.random.default_rng(seed)
    ids = np.unique(rng.integers(0, 2**64 - 1, size=n + 64, dtype=np.uint64, endpoint=True))
    while ids.size < n:
        ids = np.unique(np.concatenate([ids, rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64, endpoint=True)]))
    ids = ids[rng.permutation(ids.size)[:n]]
    return ids


def grid_graph(n, seed=None):
    """n x n grid, edges right and down; vertex ids permuted into distinct
    uint64 values from `seed` (identity ids when seed is None)."""
    idx = np.arange(n * n, dtype=np.int64).reshape(n, n)
    right = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    down = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    e = np.concatenate([right, down], axis=0)
    ids = np.arange(n * n, dtype=np.uint64) if seed is None else distinct_u64(n * n, seed)
    return with_rowid(ids[e])


def layered_dag(seed, nvert=10000, nlayers=12, nedges=14000, lam=1.4):
    """Random DAG of `nlayers` layers; out-degree Poisson(lam) towards the next
    layer, topped up or trimmed to exactly `nedges` distinct edges; uint64 ids."""
    rng = np.random.default_rng(seed)
    sizes = [nvert // nlayers + (1 if i < nvert % nlayers else 0) for i in range(nlayers)]
    starts = np.cumsum([0] + sizes)
    edges = []
    seen = set()
    for L in range(nlayers - 1):
        for v in range(starts[L], starts[L + 1]):
            k = min(int(rng.poisson(lam)), sizes[L + 1])
            for t in rng.choice(sizes[L + 1], size=k, replace=False):
                e = (v, int(starts[L + 1] + t))
                if e not in seen:
                    seen.add(e)
                    edges.append(e)
    while len(edges) < nedges:
        L = int(rng.integers(0, nlayers - 1))
        v = int(starts[L] + rng.integers(0, sizes[L]))
        t = int(starts[L + 1] + rng.integers(0, sizes[L + 1]))
        if (v, t) not in seen:
            seen.add((v, t))
            edges.append((v, t))
    edges = edges[:nedges]
    ids = distinct_u64(nvert, seed + 7919)
    e = np.asarray(edges, dtype=np.int64)
    return with_rowid(ids[e])


def random_graph(n, m, seed):
    """Uniform random directed graph with n vertices and m distinct edges, no
    self-loops, uint64 vertex ids (the batched/served TC instances); edges are
    kept in first-draw order."""
    rng = np.random.default_rng(seed)
    k = m + m // 4 + 64
    while True:
        u = rng.integers(0, n, size=k)
        v = rng.integers(0, n, size=k)
        ok = u != v
        u, v = u[ok], v[ok]
        _, first = np.unique(u.astype(np.int64) * n + v, return_index=True)
        if first.size >= m:
            break
        k *= 2
    sel = np.sort(first)[:m]
    ids = distinct_u64(n, seed + 104729)
    return with_rowid(ids[np.stack([u[sel], v[sel]], axis=1)])


def log_uniform_instances(count=256, seed=4, nmin=1000, nmax=50000, density=1.3):
    """Batched TC instances: vertex count log-uniform in [nmin, nmax], edge
    count density x vertex count, one seed per instance."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count))).astype(np.int64)
    return [(int(n), int(round(density * n)), seed * 100003 + i) for i, n in enumerate(ns)]


# ------------------------------------------------------------ closed forms ---

def grid_tc_size(n):
    return (n * (n + 1) // 2) ** 2 - n * n


def grid_N(n, d):
    """number of (source, target) pairs at grid distance d"""
    return sum((n - a) * (n - (d - a)) for a in range(max(0, d - n + 1), min(d, n - 1) + 1))


def grid_deltas(n):
    """fixed-point delta sequence of right-linear TC on the n x n grid:
    distances 2 .. 2n-2, then the empty iteration"""
    return [grid_N(n, d) for d in range(2, 2 * n - 1)] + [0]


def bfs_distance_histogram(pairs):
    """For right-linear TC seeded with the edges: the k-th fixed-point delta is
    the number of (x, z) with shortest-path length k+1 (x == z counts through
    its shortest cycle).  Returns (final |TC|, delta list ending in 0)."""
    adj = {}
    for u, v in pairs:
        adj.setdefault(u, set()).add(v)
    hist = {}
    total = 0
    for x in adj:
        dist = {}
        q = deque()
        for v in adj[x]:
            if v not in dist:
                dist[v] = 1
                q.append(v)
        while q:
            u = q.popleft()
            for v in adj.get(u, ()):
                if v not in dist:
                    dist[v] = dist[u] + 1
                    q.append(v)
        total += len(dist)
        for d in dist.values():
            hist[d] = hist.get(d, 0) + 1
    maxd = max(hist) if hist else 1
    return total, [hist.get(d, 0) for d in range(2, maxd + 1)] + [0]


def tc_size_condensation(pairs):
    """|TC| of a directed graph without enumerating paths: strongly connected
    components (scipy), then reachability sets as bitsets over the vertices,
    propagated over the condensation DAG in reverse topological order; every
    vertex of a component reaches everything its component reaches, and a
    component reaches its own members only through a cycle (size > 1 or a
    self-loop).  Returns (|TC|, number of vertices)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    verts, idx = np.unique(p, return_inverse=True)
    n = verts.size
    idx = idx.reshape(-1, 2)
    a = csr_matrix((np.ones(len(idx), dtype=np.int8), (idx[:, 0], idx[:, 1])), shape=(n, n))
    ncomp, comp = connected_components(a, directed=True, connection="strong")
    size = np.bincount(comp, minlength=ncomp)
    self_loop = np.zeros(ncomp, dtype=bool)
    self_loop[comp[idx[idx[:, 0] == idx[:, 1], 0]]] = True
    cu, cv = comp[idx[:, 0]], comp[idx[:, 1]]
    cross = cu != cv
    succ = [[] for _ in range(ncomp)]
    for u, v in set(zip(cu[cross].tolist(), cv[cross].tolist())):
        succ[u].append(v)
    # topological order of the condensation (Kahn), then reverse
    indeg = np.zeros(ncomp, dtype=np.int64)
    for u in range(ncomp):
        for v in succ[u]:
            indeg[v] += 1
    order = [u for u in range(ncomp) if indeg[u] == 0]
    for u in order:
        for v in succ[u]:
            indeg[v] -= 1
            if indeg[v] == 0:
                order.append(v)
    words = (n + 63) // 64
    members = [[] for _ in range(ncomp)]
    for v, c in enumerate(comp):
        members[c].append(v)
    own = {}

    def own_bits(c):
        if c not in own:
            b = np.zeros(words, dtype=np.uint64)
            m = np.asarray(members[c], dtype=np.int64)
            np.bitwise_or.at(b, m >> 6, np.uint64(1) << (m & 63).astype(np.uint64))
            own[c] = b
        return own[c]

    reach = {}
    total = 0
    for u in reversed(order):
        r = np.zeros(words, dtype=np.uint64)
        for v in succ[u]:
            r |= reach[v]
            if size[v] == 1:
                x = members[v][0]
                r[x >> 6] |= np.uint64(1) << np.uint64(x & 63)
            else:
                r |= own_bits(v)
        if size[u] > 1 or self_loop[u]:
            r |= own_bits(u)
        reach[u] = r
        total += int(size[u]) * int(np.unpackbits(r.view(np.uint8)).sum())
    return total, n


def tc_pairs_bruteforce(pairs):
    """set of (x, z) with a path x -> z"""
    adj = {}
    for u, v in pairs:
        adj.setdefault(u, set()).add(v)
    out = set()
    for x in adj:
        seen = set()
        stack = list(adj[x])
        while stack:
            u = stack.pop()
            if u in seen:
                continue
            seen.add(u)
            stack.extend(adj.get(u, ()))
        out.update((x, z) for z in seen)
    return out


def tc_join_pairs(pairs):
    """joined tuples of right-linear TC: sum over paths x -> y of outdeg(y)"""
    outdeg = {}
    for u, _ in pairs:
        outdeg[u] = outdeg.get(u, 0) + 1
    return sum(outdeg.get(z, 0) for (_, z) in tc_pairs_bruteforce(pairs))


def sg_bruteforce(pairs):
    """same-generation closure by naive iteration over python sets"""
    par = {}
    for p, x in pairs:
        par.setdefault(p, set()).add(x)
    sg = set()
    for kids in par.values():
        for x in kids:
            for y in kids:
                if x != y:
                    sg.add((x, y))
    while True:
        new = set()
        for (a, bb) in sg:
            for x in par.get(a, ()):
                for y in par.get(bb, ()):
                    if (x, y) not in sg:
                        new.add((x, y))
        if not new:
            return sg
        sg |= new


# ------------------------------------------------- facade program runs ------

def run_facade_program(binary, data_name, root, nb, dump_dir=None, args=(), timeout=600):
    """Run a reference program built against the facade the way its sources
    expect: cwd = <root>/tests, with <root>/data/<data_name> pointing at the
    fixture copy under tests/golden.  Returns the completed process."""
    import subprocess
    os.makedirs(os.path.join(root, "tests"), exist_ok=True)
    os.makedirs(os.path.join(root, "data"), exist_ok=True)
    link = os.path.join(root, "data", data_name)
    if not os.path.exists(link):
        os.symlink(os.path.join(GOLDEN, data_name), link)
    env = dict(os.environ, PRA_BUCKETS=str(nb))
    if dump_dir:
        os.makedirs(dump_dir, exist_ok=True)
        env["PRA_DUMP_DIR"] = str(dump_dir)
    return subprocess.run([binary, *args], cwd=os.path.join(root, "tests"), env=env, capture_output=True,
                          text=True, timeout=timeout)


def dump_digest(dump_dir):
    """{relation: [rows, sha256 of the dumped bytes]} of a PRA_DUMP_DIR dump."""
    import hashlib
    out = {}
    for f in sorted(os.listdir(dump_dir)):
        if not f.endswith("_full"):
            continue
        with open(os.path.join(dump_dir, f + ".size")) as s:
            rows = int(s.read().split()[0])
        with open(os.path.join(dump_dir, f), "rb") as d:
            out[f[:-len("_full")]] = [rows, hashlib.sha256(d.read()).hexdigest()]
    return out


def total_facts(stdout):
    """the last "Total facts across all relations N" line (lie.cpp:128-150)"""
    import re
    m = re.findall(r"Total facts across all relations (\d+)", stdout)
    return int(m[-1]) if m else None


def nonlinear_tc_program(b, edges):
    """Doubling transitive closure written with a reversed index relation:
         Tr(y, x)  <- acopy of T(x, y)              (re-keyed on y, id kept)
         T(x, z)   <- Tr(y, x) DELTA JOIN T(y, z) FULL   on y
         T(x, z)   <- Tr(y, x) FULL  JOIN T(y, z) DELTA  on y  (RIGHT join)
    T's full version is both a join inner and changes every iteration; Tr's
    full version is the outer of the RIGHT join.  Returns (T, Tr, scc)."""
    T = b.rel_create(2, 1, 410)
    Tr = b.rel_create(2, 1, 411)
    b.load_rows(T, edges, FULL)
    s = b.scc_create(True, 1)
    b.scc_add_relation(s, T, True)
    b.scc_add_relation(s, Tr, False)
    b.add_acopy(s, Tr, T, DELTA, [1, 0, 2])
    b.add_join(s, T, Tr, DELTA, T, FULL, [1, 3])
    b.add_join(s, T, Tr, FULL, T, DELTA, [1, 3])
    return T, Tr, s


def random_program(b, seed, nbase=2, nderived=4, dom=9):
    """A random two-SCC program over small relations (values in [0, dom)):
    SCC 1 (runs once) seeds the derived relations from the base ones with
    copies, filters and joins over FULL versions; SCC 2 (to a fixed point)
    applies random LEFT (DELTA x FULL) and RIGHT (FULL x DELTA) joins, copies,
    filters and acopies.  Every derived relation is the target of either acopy
    rules or join/copy/filter rules, never both; joined relations share their
    join-column count; values in data columns only come from data columns
    (ids never feed new tuples), so the fixed point is finite.  Returns the
    relation handles, the SCC handles and a description of the rules."""
    rng = random.Random(seed)
    specs = []
    for i in range(nbase + nderived):
        arity = rng.choice([1, 2, 2, 3])
        jcc = 1 if arity == 1 or rng.random() < 0.7 else 2
        specs.append((arity, jcc))
    rels = [b.rel_create(a, j, 600 + i) for i, (a, j) in enumerate(specs)]
    for i, (a, _) in enumerate(specs):
        n = rng.randint(0 if i >= nbase else 8, 40)
        tuples = sorted({tuple(rng.randrange(dom) for _ in range(a)) for _ in range(n)})
        if tuples:
            b.load_rows(rels[i], np.array([list(t) + [7000 * (i + 1) + k] for k, t in enumerate(tuples)],
                                          dtype=np.uint64), FULL)
    derived = list(range(nbase, nbase + nderived))
    acopy_target = {d: rng.random() < 0.3 for d in derived}
    rules = []

    def data_proj(a0, a1, jcc, n_out, right_cols=True):
        cols = list(range(a0)) + ([a0 + 1 + j for j in range(a1 - jcc)] if right_cols else [])
        return [rng.choice(cols) for _ in range(n_out)]

    def add_rules(scc, sources, versions):
        for d in derived:
            ad, jd = specs[d]
            if acopy_target[d]:
                cands = [r for r in sources if specs[r][0] == ad]
                if not cands or versions == "full":
                    continue
                for _ in range(rng.randint(1, 2)):
                    src = rng.choice(cands)
                    perm = list(range(ad))
                    rng.shuffle(perm)
                    b.add_acopy(scc, rels[d], rels[src], DELTA, perm + [ad])
                    rules.append(("acopy", d, src, perm))
                continue
            for _ in range(rng.randint(1, 3)):
                kind = rng.choice(["join", "join", "copy", "filter"])
                src = rng.choice(sources)
                a0, j0 = specs[src]
                if kind == "join":
                    partners = [r for r in range(len(specs)) if specs[r][1] == j0]
                    other = rng.choice(partners)
                    a1 = specs[other][0]
                    proj = data_proj(a0, a1, j0, ad)
                    if versions == "full":
                        b.add_join(scc, rels[d], rels[src], FULL, rels[other], FULL, proj)
                    elif rng.random() < 0.7:
                        b.add_join(scc, rels[d], rels[src], DELTA, rels[other], FULL, proj)
                    else:
                        b.add_join(scc, rels[d], rels[other], FULL, rels[src], DELTA,
                                   data_proj(a1, a0, j0, ad))
                    rules.append((kind, d, src, other))
                elif kind == "filter" and a0 >= 2:
                    op = rng.randrange(6)
                    b.add_copy_filter(scc, rels[d], rels[src], FULL if versions == "full" else DELTA,
                                      [rng.randrange(a0) for _ in range(ad)], op)
                    rules.append((kind, d, src, op))
                else:
                    b.add_copy(scc, rels[d], rels[src], FULL if versions == "full" else DELTA,
                               [rng.randrange(a0) for _ in range(ad)])
                    rules.append(("copy", d, src))

    s1 = b.scc_create(False, 1)
    for i in range(len(specs)):
        b.scc_add_relation(s1, rels[i], False)
    add_rules(s1, list(range(nbase)), "full")
    s2 = b.scc_create(True, 2)
    for i in range(len(specs)):
        b.scc_add_relation(s2, rels[i], i >= nbase)
    add_rules(s2, derived, "delta")
    b.add_dependency(s1, s2)
    return {"rels": rels, "sccs": [s1, s2], "rules": rules}


def id_valued_program(b, n=40, seed=3):
    """Intern ids flowing into data columns after the first SCC has already
    ranked the values it saw:
         SCC 1 (once):      X(y, x)      <- copy A(x, y) FULL
         SCC 2 (to a fixed point, A status true):
                            B(id, x | y) <- acopy A(x, y | id) DELTA   [2, 0, 1]
                            O(id, x)     <- copy B DELTA               [0, 1]
    O's candidates carry A's row ids, values no relation held when the first
    SCC ran.  Returns (O, B, X, scc1, scc2)."""
    rng = random.Random(seed)
    pairs = sorted({(rng.randrange(20), rng.randrange(20)) for _ in range(n)})
    A = b.rel_create(2, 1, 31)
    X = b.rel_create(2, 1, 32)
    B = b.rel_create(2, 1, 33)
    O = b.rel_create(2, 1, 34)
    b.load_rows(A, np.array([[x, y, 1000003 + 7 * i] for i, (x, y) in enumerate(pairs)], dtype=np.uint64), FULL)
    s1 = b.scc_create(False, 1)
    b.scc_add_relation(s1, A, False)
    b.scc_add_relation(s1, X, False)
    b.add_copy(s1, X, A, FULL, [1, 0])
    s2 = b.scc_create(True, 2)
    b.scc_add_relation(s2, A, True)
    b.scc_add_relation(s2, B, False)
    b.scc_add_relation(s2, O, False)
    b.add_acopy(s2, B, A, DELTA, [2, 0, 1])
    b.add_copy(s2, O, B, DELTA, [0, 1])
    b.add_dependency(s1, s2)
    return O, B, X, s1, s2


TWO_RULE_TAG = 5


def two_rule_program(b):
    """Two join rules into one output, one iteration, one bucket:
         O(z, x) <- A(y, x) DELTA JOIN B(y, z) FULL   (rule 0)
         O(z, x) <- A(y, x) DELTA JOIN C(y, z) FULL   (rule 1)
    A = (1,10) (2,10) (3,30); B = (1,100) (2,100) (3,300); C = (1,7) (2,7) (3,9).
    Rule 0 emits (100,10) (100,10) (300,30) and rule 1 (7,10) (7,10) (9,30), the
    second of each pair a duplicate inside its own call.  Returns O."""
    rows = lambda pairs: np.array([[k, v, 900 + i] for i, (k, v) in enumerate(pairs)], dtype=np.uint64)
    A = b.rel_create(2, 1, 1)
    B = b.rel_create(2, 1, 2)
    C = b.rel_create(2, 1, 3)
    O = b.rel_create(2, 1, TWO_RULE_TAG)
    b.load_rows(A, rows([(1, 10), (2, 10), (3, 30)]), FULL)
    b.load_rows(B, rows([(1, 100), (2, 100), (3, 300)]), FULL)
    b.load_rows(C, rows([(1, 7), (2, 7), (3, 9)]), FULL)
    s = b.scc_create(False, 1)
    b.scc_add_relation(s, A, True)
    b.scc_add_relation(s, B, False)
    b.scc_add_relation(s, C, False)
    b.scc_add_relation(s, O, False)
    b.add_join(s, O, A, DELTA, B, FULL, [3, 1])
    b.add_join(s, O, A, DELTA, C, FULL, [3, 1])
    return O


def two_rule_expected(order):
    """O's rows at one bucket when its tuples were interned in `order`
    (id = tag << 46 | bucket << 28 | seq), in download order."""
    seq = {t: i for i, t in enumerate(order)}
    return np.array([[z, x, (TWO_RULE_TAG << 46) | seq[(z, x)]] for z, x in sorted(order)], dtype=np.uint64)


# Hand-derived interning orders of two_rule_program (src/RAM/RA_tasks.cpp:312,
# 1121-1269; src/RA/parallel_join.cpp:157-230).  Without a split the arrival
# order is rule 0's outputs then rule 1's.  A call stops after the outer row
# through which its distinct outputs exceed the threshold, so with threshold 0
# every round takes one row of A from each rule: round 1 interns (100,10) then
# (7,10), round 2 re-emits them (dropped as already in newt), round 3 interns
# (300,30) then (9,30).  With threshold 1 the duplicate rows do not count, so
# each call passes 1 only at its last row: no second round, the plain order.
TWO_RULE_ORDERS = {
    None: [(100, 10), (300, 30), (7, 10), (9, 30)],
    0: [(100, 10), (7, 10), (300, 30), (9, 30)],
    1: [(100, 10), (300, 30), (7, 10), (9, 30)],
}


def run_tc_checkpointed(exe, root, nb=2, cp=1, restart_from="checkpoint-1-5"):
    """The reference TC program's checkpoint mode on data/g13236 (a checkpoint
    every `cp` loops into <root>/out), then its restart mode from
    out/<restart_from> into <root>/out2 with the final relations dumped to
    <root>/restart_dump.  Returns (out, restart_dump)."""
    import shutil
    import subprocess
    g5955 = os.path.join(root, "data", "g5955")
    os.makedirs(g5955, exist_ok=True)
    os.makedirs(os.path.join(root, "tests"), exist_ok=True)
    for suffix in ("", ".size"):
        shutil.copyfile(os.path.join(GOLDEN, "empty_path_2_1_2" + suffix), os.path.join(g5955, "path_2_1_2" + suffix))
    out = os.path.join(root, "out")
    env = dict(os.environ, PRA_BUCKETS=str(nb))
    r = subprocess.run([exe, os.path.join(GOLDEN, "g13236_edge_2_1_2"), out, str(cp)], cwd=os.path.join(root, "tests"),
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "T: {2}. (573 total facts)" in r.stdout
    dump = os.path.join(root, "restart_dump")
    os.makedirs(dump, exist_ok=True)
    r = subprocess.run([exe, "--restart", os.path.join(out, restart_from), os.path.join(root, "out2"), str(cp)],
                       cwd=os.path.join(root, "tests"), env=dict(env, PRA_DUMP_DIR=dump), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "T: {2}. (573 total facts)" in r.stdout, r.stdout[-2000:]
    return out, dump


def checkpoint_digest(out):
    """{checkpoint dir: {file: [bytes, sha256]}} of a checkpoint output directory"""
    import hashlib
    res = {}
    for d in sorted(os.listdir(out)):
        files = {}
        for f in sorted(os.listdir(os.path.join(out, d))):
            with open(os.path.join(out, d, f), "rb") as fh:
                data = fh.read()
            files[f] = [len(data), hashlib.sha256(data).hexdigest()]
        res[d] = files
    return res
