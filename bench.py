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
for p in (os.path.join(ROOT, "parallel-ra_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "joined tuples/s + fixed-point wall time, 1 MI355X; batched vs plain (bit-exact)"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["tc", "sg"], default="tc",
                    help="tc: config 2 grid transitive closure (default); sg: config 3 same generation")
    ap.add_argument("--grid", type=int, default=150)
    ap.add_argument("--buckets", type=int, default=8)
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              f"(the multi-GPU exchange is a design, not code).", file=sys.stderr)
        sys.exit(2)

    import pra_amd
    from programs import grid_graph, grid_tc_size, layered_dag, sg_program, tc_program

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

    ctxs = [prepared() for _ in range(args.steps)]
    if dom_bit is not None:
        for c, _ in ctxs:
            c.set_profiling(True, mask=1 << dom_bit)
    for c, _ in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    per = []
    for c, _ in ctxs:
        s0 = time.perf_counter()
        c.execute()
        c.synchronize()
        per.append(time.perf_counter() - s0)
    total = time.perf_counter() - t0

    joined = 0
    iters = 0
    kstats = {}
    for c, (T, G, sfix, scopy) in ctxs:
        assert c.count(T) == expect_tuples, c.count(T)
        joined += c.joined()
        iters = len(c.history(sfix))
        for k, v in c.kernel_stats().items():
            d = kstats.setdefault(k, {"launches": 0, "ms": 0.0, "bytes": 0.0})
            for f in d:
                d[f] += v[f]
        c.close()
    log("per-step s:", ", ".join(f"{x:.3f}" for x in per))

    roofline = None
    if not args.no_profile and kstats:
        dom = max(kstats, key=lambda k: kstats[k]["ms"])
        d = kstats[dom]
        avg_ms = d["ms"] / max(1, d["launches"])
        bytes_per = d["bytes"] / max(1, d["launches"])
        achieved = bytes_per / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_us": round(avg_ms * 1e3, 2), "launches_per_step": d["launches"] // args.steps,
                    "algorithmic_bytes_per_launch": round(bytes_per)}

    cpu = None if (args.no_cpu or args.workload != "tc") else cpu_baseline(edges, nb, args.cpu_iters)

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
                   "joined_tuples_per_step": joined // args.steps},
        "fixed_point_wall_s": total / args.steps,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
