"""The autotuner tries every variant of its knobs on live iterations (dense
claim table size, newt scatter width) and keeps the cheapest: a tuned solve
gives the same rows, ids, per-batch deltas and join pairs as a solve with the
fixed defaults, and the tuner reports what it measured."""
import os

import numpy as np
import pytest

from programs import grid_graph, grid_tc_size, tc_program

pytestmark = pytest.mark.gpu


def solve(n, tuned):
    import pra_amd
    if tuned:
        os.environ.pop("PRA_AUTOTUNE", None)
    else:
        os.environ["PRA_AUTOTUNE"] = "0"
    try:
        c = pra_amd.Context(nbuckets=8, batch_size=1)
    finally:
        os.environ.pop("PRA_AUTOTUNE", None)
    try:
        T, G, sfix, _ = tc_program(c, grid_graph(n, seed=n))
        c.execute()
        return c.rows(T), c.history(sfix), c.joined()
    finally:
        c.close()


def test_tuned_solve_equals_fixed_defaults():
    import pra_amd
    n = 100
    print(f"[tune] grid {n}, fixed defaults", flush=True)
    base = solve(n, tuned=False)
    assert base[0].shape[0] == grid_tc_size(n)
    for k in range(2):
        print(f"[tune] grid {n}, tuned solve {k}", flush=True)
        got = solve(n, tuned=True)
        np.testing.assert_array_equal(got[0], base[0])
        assert got[1] == base[1] and got[2] == base[2]
    rep = pra_amd.autotune_report()
    knobs = {k for _, k, _, _, _ in rep}
    assert {"dense_table_slots_log2_minus_1", "newt_scatter_wide"} <= knobs, rep
    # every band that settled measured all of its variants
    for dev, knob, band, choice, means in rep:
        if choice >= 0:
            assert len(means) == (3 if knob.startswith("dense") else 2) and 0 <= choice < len(means), (knob, band)
    assert any(choice >= 0 for _, _, _, choice, _ in rep), rep
