"""Writes serve_tc_sizes_<count>.json: |TC| of every graph in the served
request stream of random graphs (programs.log_uniform_instances with the serving
seed 5, vertex counts log-uniform in [nmin, nmax], 1.3 edges per vertex),
computed on the CPU by SCC condensation (programs.tc_size_condensation).

The serving benchmark checks every served result against these sizes; for
10,000 requests of up to 50,000 vertices the count takes about an hour on one
core, so it is computed once, in parallel, and committed.

Usage: python tests/golden/make_serve_sizes.py [--count 10000] [--procs 8]
"""
import argparse
import hashlib
import json
import os
import sys
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from programs import log_uniform_instances, random_graph, tc_size_condensation  # noqa: E402

SEED = 5


def tc_size(inst):
    n, m, sd = inst
    return int(tc_size_condensation(random_graph(n, m, sd)[:, :2])[0])


def stream_digest(inst):
    """sha256 of the (n, m, seed) list of the request stream"""
    return hashlib.sha256(json.dumps([list(map(int, i)) for i in inst]).encode()).hexdigest()


def fixture_path(count, nmin, nmax):
    return os.path.join(HERE, f"serve_tc_sizes_{count}_n{nmin}-{nmax}.json")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=10000)
    ap.add_argument("--nmin", type=int, default=1000)
    ap.add_argument("--nmax", type=int, default=50000)
    ap.add_argument("--procs", type=int, default=8)
    a = ap.parse_args()
    inst = log_uniform_instances(count=a.count, seed=SEED, nmin=a.nmin, nmax=a.nmax)
    with Pool(a.procs) as pool:
        sizes = pool.map(tc_size, inst, chunksize=8)
    out = {"count": a.count, "seed": SEED, "nmin": a.nmin, "nmax": a.nmax,
           "stream_sha256": stream_digest(inst), "tc_sizes": sizes}
    with open(fixture_path(a.count, a.nmin, a.nmax), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
