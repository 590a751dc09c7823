"""Writes tests/golden/ref_vectors.json: golden vectors printed by the
reference itself, built from its own sources into oracle/_ref (make -C oracle
ref) and run under the image's mpirun on k ranks.

Each case is a program from tests/programs.py, recorded through the same
program-building calls the oracle and the HIP binding take, then run as
  * the reference's tests/transitive_closure.cpp, unchanged (TC cases), or
  * oracle/ref_driver, a driver on the reference's C++ API (other programs),
with a checkpoint after every batch (LIE::execute, cp_iteration 1, one file per
rank: src/lie/lie.cpp:353-423, src/relation/balanced_hash_relation.cpp:109-235).
From the checkpoints and the final dump the fixture keeps, per case and rank
count: the SCCs in execution order with each batch's global delta (the
history of RAM::check_for_fixed_point), each batch's per-rank delta of every
SCC relation (rank = bucket), and every relation's final full version in the
reference's row order (rank by rank, each rank's trie order) -- the rows
themselves, intern-id column included, or their sha256 when large.

Run here (the reference tree and mpirun exist only in this container):
    python3 tests/golden/make_ref_vectors.py
"""
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import programs as P  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref")
MPIRUN = "/opt/conda/bin/mpirun"
OUT = os.path.join(HERE, "ref_vectors.json")
ROWS_INLINE = 700       # final versions up to this many rows are stored row by row


class Recorder:
    """Program-building backend that records a program (the calls of
    tests/programs.py) for the reference driver."""

    def __init__(self):
        self.rels = []      # (arity, jcc, tag)
        self.data = {}      # rel -> rows
        self.sccs = []      # (fixpoint, id)
        self.scc_rels = []  # (scc, rel, status)
        self.rules = []     # text lines
        self.deps = []
        self.logged = []    # SCCs printed after each batch (RAM::enable_logging)

    def rel_create(self, arity, jcc, tag):
        self.rels.append((arity, jcc, tag))
        return len(self.rels) - 1

    def load_rows(self, rel, rows, version=P.FULL):
        assert version == P.FULL
        self.data[rel] = np.asarray(rows, dtype=np.uint64).reshape(-1, self.rels[rel][0] + 1)

    def scc_create(self, fixpoint, sid):
        self.sccs.append((bool(fixpoint), int(sid)))
        return len(self.sccs) - 1

    def scc_add_relation(self, scc, rel, status):
        self.scc_rels.append((scc, rel, int(bool(status))))

    def add_join(self, scc, out, in0, v0, in1, v1, proj):
        self.rules.append(f"join {scc} {out} {in0} {v0} {in1} {v1} " + " ".join(map(str, proj)))

    def add_copy(self, scc, out, src, v, reorder):
        self.rules.append(f"copy {scc} {out} {src} {v} " + " ".join(map(str, reorder)))

    def add_acopy(self, scc, out, src, v, reorder):
        self.rules.append(f"acopy {scc} {out} {src} {v} " + " ".join(map(str, reorder)))

    def add_copy_filter(self, scc, out, src, v, reorder, op):
        self.rules.append(f"filter {scc} {out} {src} {v} {op} " + " ".join(map(str, reorder)))

    def add_dependency(self, before, after):
        self.deps.append((before, after))

    def name(self, rel):
        return f"R{rel}"

    def write(self, d, batch):
        lines = []
        for i, (a, j, t) in enumerate(self.rels):
            path = os.path.join(d, f"in_{i}")
            P.write_bin(path, self.data.get(i, np.zeros((0, a + 1), dtype=np.uint64)))
            lines.append(f"rel {self.name(i)} {a} {j} {t} {path}")
        for fx, sid in self.sccs:
            lines.append(f"scc {sid} {int(fx)}")
        for s, r, st in self.scc_rels:
            lines.append(f"sccrel {s} {r} {st}")
        lines += self.rules
        for a, b in self.deps:
            lines.append(f"dep {a} {b}")
        for a in self.logged:
            lines.append(f"log {a}")
        lines.append(f"batch {batch}")
        spec = os.path.join(d, "spec")
        with open(spec, "w") as f:
            f.write("\n".join(lines) + "\n")
        return spec


def run_reference(rec, k, batch, tc_edges=None):
    """Run the recorded program on k ranks; returns the parsed result."""
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "out")
        os.makedirs(out)
        if tc_edges is not None:
            # the reference TC program, unchanged: reads its empty T seed from
            # ../data/g5955/path_2_1_2 relative to its working directory
            os.makedirs(os.path.join(d, "tests"))
            os.makedirs(os.path.join(d, "data", "g5955"))
            P.write_bin(os.path.join(d, "data", "g5955", "path_2_1_2"), np.zeros((0, 3), dtype=np.uint64))
            P.write_bin(os.path.join(d, "edges"), tc_edges)
            names = {0: "T", 1: "G"}
            # tc_program creates the fixed-point SCC, then the copy SCC: the
            # reference program's RAM(true, 1) and RAM(false, 2)
            rec.sccs = [(True, 1), (False, 2)]
            cmd = [MPIRUN, "-n", str(k), os.path.join(REF_BIN, "TC"), os.path.join(d, "edges"), out, "1"]
            cwd = os.path.join(d, "tests")
            assert batch == 1
        else:
            spec = rec.write(d, batch)
            names = {i: rec.name(i) for i in range(len(rec.rels))}
            cmd = [MPIRUN, "-n", str(k), os.path.join(REF_BIN, "ref_driver"), spec, out]
            cwd = d
        r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): {r.stdout[-2000:]}{r.stderr[-2000:]}")
        order = [int(x) for x in re.findall(r"Executing SCC (-?\d+)", r.stdout)]
        return parse(out, order, names, rec, k, final_dir=None if tc_edges is not None else os.path.join(out, "final"))


def rows_of(path, w):
    if not os.path.exists(path):
        return np.zeros((0, w), dtype=np.uint64)
    a = np.fromfile(path, dtype="<u8")
    assert a.size % w == 0, path
    return a.reshape(-1, w).astype(np.uint64)


def parse(out, order, names, rec, k, final_dir):
    widths = {i: rec.rels[i][0] + 1 for i in names}
    sid_of = {i: sid for i, (_, sid) in enumerate(rec.sccs)}
    scc_members = {}
    for s, r, _ in rec.scc_rels:
        scc_members.setdefault(sid_of[s], []).append(r)
    cps = {}
    for dn in os.listdir(out):
        m = re.fullmatch(r"checkpoint-(-?\d+)-(\d+)", dn)
        if m:
            cps.setdefault(int(m.group(1)), []).append((int(m.group(2)), os.path.join(out, dn)))
    sccs = []
    last_dir = None
    for sid in order:
        batches = sorted(cps.get(sid, []))
        hist, per_rank = [], []
        for _, cp in batches:
            pr = [[rows_of(os.path.join(cp, f"{names[r]}_delta_{q}"), widths[r]).shape[0] for q in range(k)]
                  for r in scc_members[sid]]
            per_rank.append(pr)
            hist.append(int(sum(map(sum, pr))))
            last_dir = cp
        idx = [i for i, (_, x) in enumerate(rec.sccs) if x == sid]
        sccs.append({"scc": idx[0], "history": hist, "delta_per_rank": per_rank})
    final = {}
    for r in names:
        if final_dir is not None:
            rows = np.concatenate([rows_of(os.path.join(final_dir, f"{names[r]}_full_{q}"), widths[r])
                                   for q in range(k)])
        else:
            # TC: the last checkpoint follows the last batch of the last SCC
            # (delta empty at the fixed point), so its full versions are final
            rows = np.concatenate([rows_of(os.path.join(last_dir, f"{names[r]}_full_{q}"), widths[r])
                                   for q in range(k)])
        entry = {"rows_count": int(rows.shape[0]),
                 "sha256": hashlib.sha256(np.ascontiguousarray(rows, dtype="<u8").tobytes()).hexdigest()}
        if rows.shape[0] <= ROWS_INLINE:
            entry["rows"] = [[int(x) for x in row] for row in rows]
        final[str(r)] = entry
    return {"sccs": sccs, "final": final}


def tiny8_edges():
    """8 edges over 6 vertices with a cycle and converging paths: at 3 ranks
    the receiving buckets get arrivals from several source buckets"""
    return P.with_rowid([(1, 2), (2, 3), (3, 1), (2, 4), (4, 5), (5, 6), (3, 5), (6, 2)])


def two_rule_wide(b):
    """Two join rules into one output over 12 outer rows whose outputs collide
    across source buckets and across the rules:
         O(z, x) <- A(y, x) DELTA JOIN B(y, z) FULL   (rule 0)
         O(z, x) <- A(y, x) DELTA JOIN C(y, z) FULL   (rule 1)"""
    rows = lambda pairs, base: np.array([[k, v, base + i] for i, (k, v) in enumerate(pairs)], dtype=np.uint64)
    A = b.rel_create(2, 1, 11)
    B = b.rel_create(2, 1, 12)
    C = b.rel_create(2, 1, 13)
    O = b.rel_create(2, 1, 15)
    b.load_rows(A, rows([(y, 10 * (y % 4)) for y in range(1, 13)], 100), P.FULL)
    b.load_rows(B, rows([(y, 500 + y % 3) for y in range(1, 13)] + [(y, 600) for y in range(2, 13, 2)], 200), P.FULL)
    b.load_rows(C, rows([(y, 500 + (y + 1) % 3) for y in range(1, 13)] + [(y, 700) for y in range(1, 13, 3)], 300),
                P.FULL)
    s = b.scc_create(False, 1)
    b.scc_add_relation(s, A, True)
    b.scc_add_relation(s, B, False)
    b.scc_add_relation(s, C, False)
    b.scc_add_relation(s, O, False)
    b.add_join(s, O, A, P.DELTA, B, P.FULL, [3, 1])
    b.add_join(s, O, A, P.DELTA, C, P.FULL, [3, 1])
    return O


def small_dag():
    """same-generation input: 8 vertices in 3 layers"""
    return P.with_rowid([(1, 2), (1, 3), (1, 4), (2, 5), (3, 5), (3, 6), (4, 7), (4, 8), (2, 8)])


def threshold_split(b):
    """Two join rules into one output where rule 0's call emits 20,500,000
    distinct tuples in one iteration -- past the reference's 20,000,000-output
    local-join threshold (src/RAM/RA_tasks.cpp:312, src/RA/parallel_join.cpp:
    157-230) -- so the call stops after the outer row that passes it and
    execute_in_batches runs a second round: rule 0's first 4,001 rows, then
    all of rule 1, then rule 0's last 99 rows reach the interning.
         O(z, x) <- A(y, x) DELTA JOIN B(y, z) FULL   (rule 0: 4,100 x 5,000)
         O(z, x) <- A(y, x) DELTA JOIN C(y, z) FULL   (rule 1: 4,100 x 10)"""
    A = b.rel_create(2, 1, 21)
    B = b.rel_create(2, 1, 22)
    C = b.rel_create(2, 1, 23)
    O = b.rel_create(2, 1, 25)
    x = np.arange(4100, dtype=np.uint64)
    b.load_rows(A, np.stack([np.ones_like(x), x, 1000 + x], axis=1), P.FULL)
    z = np.arange(5000, dtype=np.uint64)
    b.load_rows(B, np.stack([np.ones_like(z), z, 20000 + z], axis=1), P.FULL)
    zc = np.arange(9000000, 9000010, dtype=np.uint64)
    b.load_rows(C, np.stack([np.ones_like(zc), zc, 30000 + np.arange(10, dtype=np.uint64)], axis=1), P.FULL)
    s = b.scc_create(False, 1)
    b.scc_add_relation(s, A, True)
    b.scc_add_relation(s, B, False)
    b.scc_add_relation(s, C, False)
    b.scc_add_relation(s, O, False)
    b.add_join(s, O, A, P.DELTA, B, P.FULL, [3, 1])
    b.add_join(s, O, A, P.DELTA, C, P.FULL, [3, 1])
    return O


def mixed_jcc(b):
    """Joins whose inputs have different join column counts.  The reference
    matches on input 0's count (src/RAM/RA_tasks.cpp:448), and a rank probes
    input 1 only in its own bucket (src/RA/parallel_join.cpp:180-186), so a
    match exists only where input 1's row falls in the same bucket under both
    counts:
         Q(z, x, c) <- R(y, x) DELTA JOIN W(y, z, c) FULL   (1 column against 2)
         R(b, a)    <- Q(a, b, c) DELTA JOIN E(a, b) FULL   (2 columns against 1)
         Q(b, a, c) <- W(a, b, c) FULL JOIN R(a, b) DELTA   (RIGHT: 2 against 1)
         R(b, x)    <- R(y, x) DELTA JOIN Q(y, b, c) FULL   (1 against 2, input 1 changing)
         R(z, x)    <- R(y, x) DELTA JOIN E(y, z) FULL      (equal counts: keeps R growing)
    seeded by R(x, y) <- E(y, x); 30 vertices, 55 edges, 3 colours."""
    rng = np.random.default_rng(97)
    pairs = set()
    while len(pairs) < 55:
        x, y = (int(v) for v in rng.integers(0, 30, 2))
        if x != y:
            pairs.add((x, y))
    pairs = sorted(pairs)
    trip = sorted({(x, y, int(rng.integers(0, 3))) for x, y in pairs[::2]} |
                  {(int(rng.integers(0, 30)), int(rng.integers(0, 30)), int(rng.integers(0, 3))) for _ in range(25)})
    E = b.rel_create(2, 1, 41)
    W = b.rel_create(3, 2, 42)
    R = b.rel_create(2, 1, 43)
    Q = b.rel_create(3, 2, 44)
    b.load_rows(E, np.array([[x, y, 5000 + i] for i, (x, y) in enumerate(pairs)], dtype=np.uint64), P.FULL)
    b.load_rows(W, np.array([[x, y, c, 6000 + i] for i, (x, y, c) in enumerate(trip)], dtype=np.uint64), P.FULL)
    fix = b.scc_create(True, 1)
    for rel, st in ((E, False), (W, False), (R, True), (Q, True)):
        b.scc_add_relation(fix, rel, st)
    b.add_join(fix, Q, R, P.DELTA, W, P.FULL, [3, 1, 4])
    b.add_join(fix, R, Q, P.DELTA, E, P.FULL, [1, 0])
    b.add_join(fix, Q, W, P.FULL, R, P.DELTA, [1, 0, 2])
    b.add_join(fix, R, R, P.DELTA, Q, P.FULL, [3, 1])
    b.add_join(fix, R, R, P.DELTA, E, P.FULL, [3, 1])
    seed = b.scc_create(False, 2)
    b.scc_add_relation(seed, E, False)
    b.scc_add_relation(seed, R, True)
    b.add_copy(seed, R, E, P.FULL, [1, 0])
    b.add_dependency(seed, fix)
    return R


def cases():
    """(name, builder(rec) -> None, tc edges or None, rank counts, batch sizes)"""
    out = []

    def tc(name, edges, ks):
        out.append((name, lambda b, e=edges: P.tc_program(b, e), edges, ks, [1]))

    tc("tc_g5955", P.golden("g5955_edge_2_1_2"), [1, 3])
    tc("tc_g13236", P.golden("g13236_edge_2_1_2"), [1, 2, 3, 5])
    tc("tc_tiny8", tiny8_edges(), [1, 2, 3])
    for seed in range(3):
        tc(f"tc_small_s{seed}", P.small_random_graph(seed), [2, 3, 5])
    tc("tc_grid30_u64", P.grid_graph(30, seed=30), [8])
    out.append(("two_rule", lambda b: P.two_rule_program(b), None, [1, 2, 3], [1]))
    out.append(("two_rule_wide", two_rule_wide, None, [1, 2, 3, 5], [1]))
    out.append(("sg_dag8", lambda b: P.sg_program(b, small_dag()), None, [1, 3], [1, 2]))
    out.append(("sg_dag600", lambda b: P.sg_program(b, P.layered_dag(11, nvert=600, nlayers=6, nedges=840)), None,
                [1, 3], [1, 3]))
    out.append(("nonlinear_tc_small_s0", lambda b: P.nonlinear_tc_program(b, P.small_random_graph(0)), None, [1, 3],
                [1]))
    for seed in range(6):
        out.append((f"random_program_s{seed}", lambda b, s=seed: P.random_program(b, s), None, [1, 3], [1, 3]))
    out.append(("id_valued", lambda b: P.id_valued_program(b), None, [1, 3], [1]))
    out.append(("threshold_split", threshold_split, None, [1], [1]))
    out.append(("mixed_jcc", mixed_jcc, None, [1, 2, 3, 5], [1, 2]))
    return out


LOG_LINE = re.compile(r"^(\S+ \d+ (FULL|DELTA) Rows \d+ columns \d+|F \[\d+\] [\d ]+|D [\d ]+)$")


def relation_dump_lines(stdout):
    """the lines relation::print writes (src/relation/balanced_hash_relation.cpp:433-522)"""
    return [l.rstrip() for l in stdout.splitlines() if LOG_LINE.match(l.rstrip())]


def write_logging_goldens():
    """RAM::enable_logging on the fixed-point SCC of the TC program, at one
    rank: the relation dumps the reference prints after every batch"""
    for name in ("g5955_edge_2_1_2", "tiny8"):
        edges = P.golden(name) if name != "tiny8" else tiny8_edges()
        rec = Recorder()
        P.tc_program(rec, edges)
        rec.sccs = [(True, 1), (False, 2)]
        rec.rels = [(2, 1, 257), (2, 1, 256)]
        rec.logged = [0]
        with tempfile.TemporaryDirectory() as d:
            spec = rec.write(d, 1)
            spec_text = open(spec).read().replace("rel R0 ", "rel T ").replace("rel R1 ", "rel G ")
            open(spec, "w").write(spec_text)
            os.makedirs(os.path.join(d, "out"))
            r = subprocess.run([MPIRUN, "-n", "1", os.path.join(REF_BIN, "ref_driver"), spec, os.path.join(d, "out")],
                               cwd=d, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
        path = os.path.join(HERE, f"ref_log_tc_{name.split('_')[0]}_k1.txt")
        with open(path, "w") as f:
            f.write("\n".join(relation_dump_lines(r.stdout)) + "\n")
        print("wrote", path)


def main():
    """all cases, or with --only NAME[,NAME..] those cases, updated in place"""
    if not os.path.exists(os.path.join(REF_BIN, "ref_driver")):
        sys.exit("build the reference checkers first: make -C oracle ref")
    only = None
    if len(sys.argv) > 2 and sys.argv[1] == "--only":
        only = set(sys.argv[2].split(","))
    result = {}
    if only is None:
        write_logging_goldens()
    else:
        with open(OUT) as f:
            result = json.load(f)
    for name, build, edges, ks, batches in cases():
        if only is not None and name not in only:
            continue
        rec = Recorder()
        build(rec)
        for k in ks:
            for bs in batches:
                key = f"{name}/k{k}/b{bs}"
                result[key] = run_reference(rec, k, bs, tc_edges=edges)
                print(key, [s["history"][:6] for s in result[key]["sccs"]], flush=True)
    with open(OUT, "w") as f:
        json.dump(result, f, separators=(",", ":"))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
