"""pra_partition_rows_noalloc (the allocation-free variant, static scratch)
against pra_partition_rows on the same inputs, bit-exact, through the C ABI:
two calls in a row, the second with a smaller input, so the second reads
nothing the first left in the shared scratch; plus the size limit."""
import numpy as np
import pytest

from partition_expect import partition_expect
from test_gpu_partition import run_partition

pytestmark = pytest.mark.gpu


def test_noalloc_matches_partition_twice():
    import pra_amd
    ctx = pra_amd.Context(nbuckets=8)
    try:
        for n, w, jcc, seed in [(1500000, 3, 1, 1), (70001, 4, 2, 2)]:
            print("case", n, w, jcc, flush=True)
            rng = np.random.default_rng(seed)
            rows = rng.integers(0, 2**64 - 1, size=(n, w), dtype=np.uint64, endpoint=True)
            rows[: n // 4, :jcc] = rows[n // 4: n // 2, :jcc]
            a_rows, a_bnd = run_partition(ctx, rows, w, jcc, noalloc=False)
            b_rows, b_bnd = run_partition(ctx, rows, w, jcc, noalloc=True)
            np.testing.assert_array_equal(b_bnd, a_bnd)
            np.testing.assert_array_equal(b_rows, a_rows)
            want, wbnd = partition_expect(rows, jcc, 8)
            np.testing.assert_array_equal(b_bnd, wbnd)
            np.testing.assert_array_equal(b_rows, want)
    finally:
        ctx.close()


def test_noalloc_largest_input_and_limit():
    import pra_amd
    ctx = pra_amd.Context(nbuckets=256)
    try:
        n = 1 << 23
        rng = np.random.default_rng(3)
        rows = rng.integers(0, 2**64 - 1, size=(n, 2), dtype=np.uint64, endpoint=True)
        b_rows, b_bnd = run_partition(ctx, rows, 2, 1, noalloc=True)
        want, wbnd = partition_expect(rows, 1, 256)
        np.testing.assert_array_equal(b_bnd, wbnd)
        np.testing.assert_array_equal(b_rows, want)
        with pytest.raises(pra_amd.PraError, match="at most"):
            ctx.partition_rows_noalloc(0, n + 1, 2, 1, 0, 0)   # null buffers too: no launch either way
    finally:
        ctx.close()
