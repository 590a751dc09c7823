"""Random multi-rule programs (tests/programs.py::random_program) on the GPU
through the C ABI against the oracle: every relation's rows, ids included,
every SCC's per-batch delta counts and the join pair count must be equal.

The programs mix LEFT and RIGHT joins, copies, filters and acopies over
relations of arity 1-3 with one or two join columns, so they reach rule
combinations that the fixed programs of the other tests do not."""
import numpy as np
import pytest

import oracle as O
from programs import random_program

pytestmark = pytest.mark.gpu


def run_both(seed, nb, batch, threshold=None):
    import pra_amd
    dom = 9 if seed % 2 else 24
    o = O.Oracle(nb)
    g = pra_amd.Context(nbuckets=nb, batch_size=batch)
    if threshold is not None:
        o.set_join_threshold(threshold)
        g.set_join_threshold(threshold)
    ho = random_program(o, seed, dom=dom)
    o.execute(batch)
    hg = random_program(g, seed, dom=dom)
    g.execute()
    return o, ho, g, hg


@pytest.mark.parametrize("nb,batch,threshold", [(1, 1, None), (3, 1, None), (3, 4, None), (8, 2, None),
                                                (1, 1, 7), (3, 2, 30)])
def test_random_programs_match_the_oracle(nb, batch, threshold):
    # a small local-join threshold splits the join calls into rounds
    for seed in range(100):
        print(f"[random program] seed={seed} nb={nb} batch={batch} threshold={threshold}", flush=True)
        o, ho, g, hg = run_both(seed, nb, batch, threshold)
        for i, (ro, rg) in enumerate(zip(ho["rels"], hg["rels"])):
            a, b = o.rows(ro), g.rows(rg)
            assert a.shape == b.shape and np.array_equal(a, b), (seed, i, a.shape, b.shape, ho["rules"])
        for so, sg in zip(ho["sccs"], hg["sccs"]):
            assert o.history(so) == g.history(sg), (seed, ho["rules"])
        assert o.joined() == g.joined(), seed
        g.close()
