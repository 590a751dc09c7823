"""Benchmark: transitive closure of the 150x150 grid (44,700 uint64 edges ->
128,233,125 tuples over 298 fixed-point iterations) on one MI355X, laid out as
the reference's 8-rank run (8 virtual buckets), reported as joined tuples/s.

A step is one whole fixed-point solve (the copy SCC and the 298-iteration join
SCC of tests/transitive_closure.cpp) with the edge relation already resident
in HBM.  Prints one JSON line; everything else goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "parallel-ra_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"),
          os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "joined tuples/s + fixed-point wall time, 1 MI355X; batched vs plain (bit-exact)"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


REF_TC = os.path.join(ROOT, "oracle", "_ref", "TC")
MPIRUN = "/opt/conda/bin/mpirun"
REF_RANKS = 8
REF_SAMPLE_GRID = 100
# the reference's own 8-rank run of the full 150x150 grid, measured in the
# build container: 77.5 s over 298 iterations
REF_FULL_GRID150_S = 77.5


def grid_join_pairs(n):
    """join pairs of semi-naive TC on the n x n grid (edges right and down):
    every path x -> y enters delta once and is joined with y's out-edges, and
    (i, j) is reached from the (i+1)(j+1)-1 cells up and to its left"""
    return sum(((i + 1) * (j + 1) - 1) * ((i < n - 1) + (j < n - 1)) for i in range(n) for j in range(n))


def cpu_baseline_reference(nb):
    """The reference itself (tests/transitive_closure.cpp built from its own
    sources into oracle/_ref/TC) under mpirun on REF_RANKS host processes, on a
    bounded sample of the same workload: the REF_SAMPLE_GRID-square grid with
    randomly permuted u64 ids.  Time = the sum of the reference's own
    per-iteration Total_time column over its fixed-point SCC.  None when the
    reference build or mpirun is absent."""
    import subprocess
    import tempfile
    import numpy as np
    from programs import grid_graph, write_bin
    if not (os.path.exists(REF_TC) and os.path.exists(MPIRUN)) or nb != REF_RANKS:
        return None
    n = REF_SAMPLE_GRID
    with tempfile.TemporaryDirectory() as d:
        # the program reads its (empty) T seed from ../data/g5955/path_2_1_2
        os.makedirs(os.path.join(d, "tests"))
        os.makedirs(os.path.join(d, "data", "g5955"))
        write_bin(os.path.join(d, "data", "g5955", "path_2_1_2"), np.zeros((0, 3), dtype=np.uint64))
        write_bin(os.path.join(d, "edges"), grid_graph(n, seed=150))
        r = subprocess.run([MPIRUN, "-n", str(REF_RANKS), REF_TC, os.path.join(d, "edges")], cwd=os.path.join(d, "tests"),
                           capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"reference baseline failed ({r.returncode}): {r.stderr[-500:]}")
        return None
    total, iters, in_fix = 0.0, 0, False
    for line in r.stdout.splitlines():
        if line.startswith("-------------------Executing SCC"):
            in_fix = line.split()[-1].startswith("1-")   # the fixed-point SCC is RAM(true, 1)
        f = line.split("\t")
        if in_fix and len(f) == 9 and f[0].isdigit():
            total += float(f[8])
            iters += 1
    facts = (n * (n + 1) // 2) ** 2 - n * n
    if f"({facts} total facts)" not in r.stdout or not iters:
        log("reference baseline: unexpected output")
        return None
    j = grid_join_pairs(n)
    return {"value": j / total, "unit": "joined tuples/s", "cores": REF_RANKS, "kind": "reference",
            "sample": f"the reference's tests/transitive_closure.cpp built from its sources (oracle/_ref/TC), "
                      f"mpirun -n {REF_RANKS}, on the {n}x{n} grid with permuted u64 ids: {j} joined tuples, "
                      f"{facts} tuples, {iters} iterations in {total:.2f} s (sum of its per-iteration Total_time); "
                      f"its full 150x150 run at {REF_RANKS} ranks took "
                      f"{REF_FULL_GRID150_S} s ({grid_join_pairs(150) / REF_FULL_GRID150_S / 1e6:.2f} M joined tuples/s)"}


def cpu_baseline(edges, nb, iters):
    """The oracle (a scalar C restatement of the reference loop) on the first
    `iters` fixed-point iterations of the same workload."""
    import oracle as O
    from programs import tc_program
    o = O.Oracle(nb)
    o.set_max_iterations(iters + 1)          # + the seeding copy iteration
    T, G, sfix, _ = tc_program(o, edges)
    t0 = time.perf_counter()
    o.execute(1)
    dt = time.perf_counter() - t0
    j = o.joined()
    return {"value": j / dt, "unit": "joined tuples/s", "cores": 1, "kind": "port",
            "sample": f"grid150 TC, {nb} buckets, first {iters} fixed-point iterations: "
                      f"{j} joined tuples in {dt:.2f} s"}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload`, from the calibrated
    FETCH_SIZE / WRITE_SIZE counter runs recorded in profiles/pmc_traffic.json
    (None when that workload / kernel was not profiled)."""
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
            return json.load(f)[workload][kernel]["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["tc", "sg", "batch", "serve"], default="tc",
                    help="tc: transitive closure of the 150x150 grid (default); sg: same generation on a "
                         "12-layer 10K-vertex DAG; batch: 256 random graphs, batched vs plain; serve: a Poisson "
                         "request stream of random-graph reachability solves")
    ap.add_argument("--count", type=int, default=256, help="batch/serve: number of instances / requests")
    ap.add_argument("--nmin", type=int, default=1000)
    ap.add_argument("--nmax", type=int, default=50000)
    ap.add_argument("--slots", type=int, default=None,
                    help="batch/serve: instances / requests in flight at most (default 256 / 64)")
    ap.add_argument("--budget-gb", type=float, default=160.0, help="batch/serve: HBM admission budget")
    ap.add_argument("--alone-n", type=int, default=10000,
                    help="batch: instances of at least this many vertices run one at a time (0: no such rule)")
    ap.add_argument("--rates", default="50,100,200,500,1000,2000", help="serve: Poisson arrival rates (req/s)")
    ap.add_argument("--grid", type=int, default=150)
    ap.add_argument("--buckets", type=int, default=8)
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-closed-form", action="store_true", help="batch/serve: skip the CPU |TC| check")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              f"(the multi-GPU exchange is a design, not code).", file=sys.stderr)
        sys.exit(2)

    import pra_amd
    from programs import grid_graph, grid_tc_size, layered_dag, sg_program, tc_program
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)

    n, nb = args.grid, args.buckets
    if args.workload == "tc":
        edges = grid_graph(n, seed=150)
        expect_tuples = grid_tc_size(n)
        workload = f"tc_grid{n}_u64ids"

        def build(c):
            T, G, sfix, scopy = tc_program(c, edges)
            return T, G, sfix, scopy
    else:
        edges = layered_dag(3)
        expect_tuples = 1392703
        workload = "same_generation_dag10k_l12"

        def build(c):
            h = sg_program(c, edges)
            return h["SG"], h["E"], h["rec"], h["base"]

    def prepared():
        c = pra_amd.Context(nbuckets=nb, device=0, batch_size=1)
        h = build(c)
        c.synchronize()
        return c, h

    # warmup solves; the last one times every kernel class to find the
    # dominant kernel, which alone is then timed live in the measured steps
    dom_bit = None
    for w in range(max(args.warmup, 0 if args.no_profile else 1)):
        c, (T, G, sfix, scopy) = prepared()
        profiled = not args.no_profile and w == max(args.warmup, 1) - 1
        if profiled:
            c.set_profiling(True)
        t0 = time.perf_counter()
        c.execute()
        c.synchronize()
        log(f"warmup {w}: {time.perf_counter() - t0:.3f} s, |T|={c.count(T)}" + (" (all kernels timed)" if profiled else ""))
        assert c.count(T) == expect_tuples
        if profiled:
            ks = c.kernel_stats()
            names = list(ks.keys())
            for k in sorted(ks, key=lambda k: -ks[k]["ms"]):
                v = ks[k]
                if v["launches"]:
                    gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0
                    log(f"  {k:20s} launches={v['launches']:7d} ms={v['ms']:9.2f} GB/s={gbs:8.1f}")
            dom_bit = names.index(max(ks, key=lambda k: ks[k]["ms"]))
        c.close()

    # measured steps: each builds its context (edges resident in HBM) outside
    # the timed region, times execute + synchronize, checks and closes it
    # before the next step, so only one solved instance is held at a time
    per = []
    joined = 0
    iters = 0
    kstats = {}
    peak_bytes = 0
    for _ in range(args.steps):
        c, (T, G, sfix, scopy) = prepared()
        if dom_bit is not None:
            c.set_profiling(True, mask=1 << dom_bit)
        c.synchronize()
        s0 = time.perf_counter()
        c.execute()
        c.synchronize()
        per.append(time.perf_counter() - s0)
        assert c.count(T) == expect_tuples, c.count(T)
        joined += c.joined()
        iters = len(c.history(sfix))
        peak_bytes = max(peak_bytes, c.device_bytes()[1])
        for k, v in c.kernel_stats().items():
            d = kstats.setdefault(k, {"launches": 0, "ms": 0.0, "bytes": 0.0})
            for f in d:
                d[f] += v[f]
        c.close()
    total = sum(per)
    log("per-step s:", ", ".join(f"{x:.3f}" for x in per))

    roofline = None
    if not args.no_profile and kstats:
        dom = max(kstats, key=lambda k: kstats[k]["ms"])
        d = kstats[dom]
        avg_ms = d["ms"] / max(1, d["launches"])
        bytes_per = d["bytes"] / max(1, d["launches"])
        achieved = bytes_per / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(workload, dom),
                    "avg_launch_us": round(avg_ms * 1e3, 2), "launches_per_step": d["launches"] // args.steps,
                    "algorithmic_bytes_per_launch": round(bytes_per)}

    cpu = None
    if not args.no_cpu and args.workload == "tc":
        cpu = cpu_baseline_reference(nb) if n == 150 else None
        if cpu is None:
            cpu = cpu_baseline(edges, nb, args.cpu_iters)

    out = {
        "metric": METRIC,
        "value": joined / total,
        "unit": "joined tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": workload, "buckets": nb, "edges": int(len(edges)),
                   "tuples": expect_tuples, "fixed_point_iterations": iters,
                   "joined_tuples_per_step": joined // args.steps, "peak_device_bytes": peak_bytes},
        "fixed_point_wall_s": total / args.steps,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": [{"knob": k, "band_log2": b, "choice": ch, "ms_per_candidate": m}
                     for _, k, b, ch, m in pra_amd.autotune_report()],
    }
    print(json.dumps(out), flush=True)


def est_bytes(n):
    """device bytes a TC instance on n vertices with m = 1.3 n edges needs at
    its peak: ~0.18 n^2 paths at up to 90 B each (the full, delta and newt
    versions in blocks up to 1.5x their size, the growth copy, the join
    candidates and the dense membership bitmap) and a scratch allowance"""
    return int(0.18 * n * n * 90) + (64 << 20)


def bench_batch(args):
    """Independent TC instances on random graphs, plain (one context per instance,
    one after another) against batched (one context whose instances run in
    lock step: every phase of an iteration is one batched launch over all of
    them; instances are queued largest first and admitted between iterations
    while their estimated bytes fit the HBM budget, up to --slots at once, and
    a finished instance is read and released at once so that the next one
    takes its place).  An instance of --alone-n vertices or more fills the
    device by itself: two such run together take longer than one after the
    other (their membership bitmaps, 12-300 MB each, evict each other from
    the 256 MB last-level cache), so they have a queue of their own and run
    one at a time, while the smaller instances ride along in the same
    lock-step launches.  Every instance is compared bit for bit (device checksum
    of its rows in order, count, join pairs) and its |TC| against the
    SCC-condensation count computed on the CPU."""
    import pra_amd
    from programs import log_uniform_instances, random_graph, tc_program, tc_size_condensation
    inst = log_uniform_instances(count=args.count, nmin=args.nmin, nmax=args.nmax)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    nb = args.buckets
    log(f"batch: {len(graphs)} instances, n in [{min(n for n, _, _ in inst)}, {max(n for n, _, _ in inst)}], "
        f"{sum(m for _, m, _ in inst)} edges")

    # plain: one instance at a time, each on its own context
    plain = []
    t_plain = 0.0
    for k, e in enumerate(graphs):
        c = pra_amd.Context(nbuckets=nb, batch_size=1)
        T, G, sfix, _ = tc_program(c, e)
        c.synchronize()
        t0 = time.perf_counter()
        c.execute()
        c.synchronize()
        t_plain += time.perf_counter() - t0
        plain.append((c.count(T), c.joined(), c.checksum(T)))
        c.close()
        if (k + 1) % 32 == 0:
            log(f"  plain: {k + 1}/{len(graphs)} instances, {t_plain:.2f} s")

    # batched: continuous admission into one context.  As in plain, the timed
    # figure starts with the inputs resident in HBM and ends before the result
    # checks: building an admitted instance's program (its edges uploaded) and
    # the checks of a finished one (its rows sorted into the reference's order
    # and checksummed) run between the steps on a synchronized stream, and
    # their time is taken out.  A finished instance is released at once while
    # others wait for memory (that release is part of the run); once the queue
    # is empty, finished instances are checked and released after the run, as
    # plain closes its contexts after its timed region
    from collections import deque
    budget = int(args.budget_gb * (1 << 30))
    slots = args.slots or 256
    order = sorted(range(len(graphs)), key=lambda i: -inst[i][0])
    def large(i):
        return args.alone_n > 0 and inst[i][0] >= args.alone_n
    big_queue = deque(i for i in order if large(i))
    queue = deque(i for i in order if not large(i))
    big_in_flight = 0
    batched = [None] * len(graphs)
    c = pra_amd.Context(nbuckets=nb, batch_size=1)
    free = list(range(slots - 1, -1, -1))
    flight = {}                        # instance id -> (graph index, handles)
    held = 0                           # estimated bytes of the instances in flight
    steps = top = 0
    t_checks = t_admit = 0.0
    finished = []                      # (instance id, graph index, handles) left for after the run
    t0 = time.perf_counter()
    def admissible():
        """the queue the next admission comes from, or None"""
        if not free:
            return None
        for q in (big_queue, queue):
            if q and (q is queue or not big_in_flight) and (not flight or held + est_bytes(inst[q[0]][0]) <= budget):
                return q
        return None

    while queue or big_queue or flight:
        if admissible():
            c.synchronize()
            a0 = time.perf_counter()
            while (q := admissible()) is not None:
                i = q.popleft()
                g = free.pop()
                c.set_group(g)
                flight[g] = (i, tc_program(c, graphs[i]))
                held += est_bytes(inst[i][0])
                big_in_flight += large(i)
            c.synchronize()
            t_admit += time.perf_counter() - a0
        top = max(top, len(flight))
        c.step()
        steps += 1
        for g in list(flight):
            if c.group_done(g):
                i, hs = flight.pop(g)
                big_in_flight -= large(i)
                if not queue and not big_queue:
                    # nothing waits for its memory: checked and released after the run
                    finished.append((g, i, hs))
                    continue
                c.synchronize()
                k0 = time.perf_counter()
                batched[i] = (c.count(hs[0]), c.group_joined(g), c.checksum(hs[0]))
                t_checks += time.perf_counter() - k0
                c.release_group(g)
                free.append(g)
                held -= est_bytes(inst[i][0])
                done = sum(b is not None for b in batched)
                if done % 32 == 0:
                    log(f"  batched: {done}/{len(graphs)} instances, {time.perf_counter() - t0:.2f} s")
    c.synchronize()
    t_batch = time.perf_counter() - t0 - t_checks - t_admit
    for g, i, hs in finished:
        batched[i] = (c.count(hs[0]), c.group_joined(g), c.checksum(hs[0]))
    peak = c.device_bytes()[1]
    c.close()
    log(f"  batched: {steps} lock-step iterations, at most {top} instances in flight, peak {peak / 2**30:.1f} GiB, "
        f"{t_admit:.2f} s of program builds and {t_checks:.2f} s of result checks taken out")
    exact = all(a == b for a, b in zip(plain, batched))
    closed = None
    if not args.no_closed_form:
        closed = True
        for i in range(len(graphs)):
            closed = closed and plain[i][0] == tc_size_condensation(graphs[i][:, :2])[0]
            if (i + 1) % 32 == 0:
                log(f"  closed form: {i + 1}/{len(graphs)} instances checked, all equal so far: {closed}")
    joined = sum(j for _, j, _ in plain)
    log(f"plain {t_plain:.3f} s, batched {t_batch:.3f} s, bit-exact per instance: {exact}, "
        f"|TC| = closed form: {closed}")
    out = {
        "metric": METRIC, "value": joined / t_batch, "unit": "joined tuples/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": 1e3 * t_batch, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"tc_batch{len(graphs)}_n{args.nmin}-{args.nmax}", "buckets": nb,
                   "budget_gb": args.budget_gb, "slots": slots, "alone_n": args.alone_n,
                   "lockstep_iterations": steps,
                   "max_in_flight": top, "tuples": sum(t for t, _, _ in plain),
                   "joined_tuples": joined},
        "plain_value": joined / t_plain, "plain_s": t_plain, "batched_s": t_batch,
        "batched_builds_s": t_admit, "batched_checks_s": t_checks,
        "batched_over_plain": t_plain / t_batch, "bit_exact_per_instance": exact, "tc_equals_closed_form": closed,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Served reachability: Poisson request streams of TC solves (sizes from the batched
    distribution) into the serving loop (continuous batching: up to --slots
    requests in flight advance together, admitted by measured device bytes
    under the HBM budget); p50/p99 latency and aggregate joined tuples/s per
    arrival rate; every result equal across rates and to the SCC-condensation
    |TC| computed on the CPU."""
    import numpy as np
    import pra_amd
    from programs import log_uniform_instances, random_graph, tc_size_condensation
    import make_serve_sizes as S
    inst = log_uniform_instances(count=args.count, seed=S.SEED, nmin=args.nmin, nmax=args.nmax)
    graphs = [random_graph(n, m, sd) for n, m, sd in inst]
    want = None
    if not args.no_closed_form:
        fixture = S.fixture_path(args.count, args.nmin, args.nmax)
        if os.path.exists(fixture):
            with open(fixture) as f:
                fx = json.load(f)
            assert fx["stream_sha256"] == S.stream_digest(inst), "stream differs from the fixture's"
            want = fx["tc_sizes"]
            log(f"closed-form |TC| of the {len(want)} requests from {os.path.relpath(fixture, ROOT)}")
        else:
            from multiprocessing import Pool
            with Pool(min(16, os.cpu_count() or 1)) as pool:
                want = pool.map(S.tc_size, inst, chunksize=4)
    budget = int(args.budget_gb * (1 << 30))
    ref = None
    rows = []
    for rate in [float(r) for r in args.rates.split(",")]:
        rng = np.random.default_rng(int(rate))
        gaps = rng.exponential(1.0 / rate, size=len(graphs))
        srv = pra_amd.Server(nbuckets=args.buckets, slots=args.slots or 64, hbm_budget=budget)
        t0 = time.perf_counter()
        tickets = []
        due = 0.0
        for k, (e, (n, _, _), gap) in enumerate(zip(graphs, inst, gaps)):
            due += gap
            now = time.perf_counter() - t0
            if due > now:
                time.sleep(due - now)
            tickets.append(srv.submit(e, est_bytes=est_bytes(n)))
            if (k + 1) % 1000 == 0:
                log(f"  rate {rate:g}: {k + 1} submitted at {time.perf_counter() - t0:.1f} s")
        res = []
        for t in tickets:
            res.append(srv.wait(t))
            if len(res) % 1000 == 0:
                log(f"  rate {rate:g}: {len(res)} served at {time.perf_counter() - t0:.1f} s")
        srv.close()
        bad = [(i, inst[i][0], r["status"], round(r["t_start"], 3), round(r["t_done"], 3))
               for i, r in enumerate(res) if r["status"] != 0]
        if bad:
            log(f"rate {rate}: {len(bad)} of {len(res)} requests failed; first (index, n, status, start, done): "
                f"{bad[:8]}")
        assert not bad
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        joined = sum(r["joined"] for r in res)
        got = [(r["tuples"], r["joined"]) for r in res]
        if ref is None:
            ref = got
        rows.append({"rate": rate, "restarts": int(sum(r["restarts"] for r in res)),
                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3), "joined_tuples_per_s": joined / span,
                     "results_equal_across_rates": got == ref,
                     "tc_equals_closed_form": None if want is None else [t for t, _ in got] == want})
        log(f"rate {rate:7.1f}/s: p50 {rows[-1]['p50_ms']:.1f} ms p99 {rows[-1]['p99_ms']:.1f} ms "
            f"{joined / span / 1e9:.3f} G joined/s, {rows[-1]['restarts']} restarts")
    out = {
        "metric": METRIC, "value": max(r["joined_tuples_per_s"] for r in rows), "unit": "joined tuples/s",
        "n_gpus": 1, "steps": len(rows), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"tc_serve{len(graphs)}_n{args.nmin}-{args.nmax}", "buckets": args.buckets,
                   "slots": args.slots or 64},
        "serve": rows, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
