"""Dense membership's claim table under every epoch width: no epoch bits
(each winner empties its slot), one bit (the table filled with zeros before
every iteration) and two bits (filled every third iteration, the slots of
the two iterations before counting as empty).  PRA_DENSE_EPOCH_BITS caps the
width; by default it is what the rank space leaves (7 bits for the 150x150
grid, whose full-size solve wraps twice).  Every width must equal the
oracle, ids and per-batch deltas included."""
import numpy as np
import pytest

import oracle as O
from programs import grid_graph, grid_tc_size, tc_program
from test_gpu_segsort import BUILDERS, Rec, context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", ["0", "1", "2"])
@pytest.mark.parametrize("k", [1, 3])
@pytest.mark.parametrize("name,build", BUILDERS)
def test_epoch_widths_equal_oracle(name, build, k, bits):
    print(f"[epochs] {name} k={k} bits={bits}", flush=True)
    g = context(k, {"PRA_DENSE_EPOCH_BITS": bits})
    try:
        rg = Rec(g)
        build(rg)
        g.execute()
        o = O.Oracle(k)
        ro = Rec(o)
        build(ro)
        o.execute(1)
        for a, b in zip(rg.rels, ro.rels):
            np.testing.assert_array_equal(g.rows(a), o.rows(b))
        for a, b in zip(rg.sccs, ro.sccs):
            assert g.history(a) == o.history(b)
        assert g.joined() == o.joined()
    finally:
        g.close()


def test_grid_epoch_widths_agree():
    """a 60x60 grid (119 iterations) gives the same rows with one epoch bit,
    with none and with the default width"""
    n = 60
    e = grid_graph(n)
    out = {}
    for bits in ("1", "0", "16"):
        print(f"[epochs] grid{n} bits={bits}", flush=True)
        g = context(4, {"PRA_DENSE_EPOCH_BITS": bits})
        try:
            T, G, sfix, _ = tc_program(g, e)
            g.execute()
            out[bits] = (g.count(T), g.checksum(T), g.history(sfix), g.joined())
        finally:
            g.close()
    assert out["1"] == out["0"] == out["16"]
    assert out["16"][0] == grid_tc_size(n)
