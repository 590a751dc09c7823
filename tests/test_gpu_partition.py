"""pra_partition_rows (the on-device routing of rows to buckets) against the
numpy restatement in tests/partition_expect.py, bit-exact, through the C ABI,
with device buffers from the HIP runtime (tests/hipmem.py)."""
import numpy as np
import pytest

from partition_expect import fnv_rows, partition_expect

pytestmark = pytest.mark.gpu


def run_partition(ctx, rows, w, jcc, noalloc=False):
    """partition rows on the device through the C ABI; returns (rows, bounds)"""
    from hipmem import DeviceBuffer
    n = rows.shape[0]
    d_rows = DeviceBuffer.from_array(rows.reshape(-1).astype(np.uint64))
    d_out = DeviceBuffer(max(1, n * w) * 8)
    d_bnd = DeviceBuffer((ctx.nb + 1) * 8)
    d_out.fill(0xAB)   # every row and bound must be written by the call
    d_bnd.fill(0xAB)
    f = ctx.partition_rows_noalloc if noalloc else ctx.partition_rows
    f(d_rows.ptr, n, w, jcc, d_out.ptr, d_bnd.ptr)
    ctx.synchronize()
    out = d_out.to_array(count=n * w).reshape(n, w)
    bnd = d_bnd.to_array(count=ctx.nb + 1)
    for b in (d_rows, d_out, d_bnd):
        b.close()
    return out, bnd


def test_fnv_restatement_known_answers():
    # textbook FNV-1a-64 vectors (also the oracle's pins)
    rows = np.array([[0, 0], [1, 0], [2**64 - 1, 0], [1, 2], [2, 1]], dtype=np.uint64)
    assert list(fnv_rows(rows[:3], 1)) == [12161962213042174405, 9929646806074584996, 10157053723145373757]
    assert list(fnv_rows(rows[3:], 2)) == [8581494755304202342, 513837244993915590]


@pytest.mark.parametrize("nb,n,w,jcc", [(1, 1000, 3, 1), (8, 0, 3, 1), (8, 1, 3, 1), (3, 2047, 3, 1),
                                        (3, 2049, 3, 1), (8, 300000, 3, 1), (5, 70000, 5, 2), (256, 500000, 2, 1),
                                        (7, 4096, 4, 0), (16, 3000000, 3, 1)])
def test_partition_matches_expectation(nb, n, w, jcc):
    import pra_amd
    print("case", nb, n, w, jcc, flush=True)
    rng = np.random.default_rng(n + 31 * nb + w)
    rows = rng.integers(0, 2**64 - 1, size=(n, w), dtype=np.uint64, endpoint=True)
    if n > 10:
        rows[: n // 3, :jcc] = rows[n // 3: 2 * (n // 3), :jcc]   # repeated keys
    ctx = pra_amd.Context(nbuckets=nb)
    try:
        out, bnd = run_partition(ctx, rows, w, jcc)
    finally:
        ctx.close()
    want, wbnd = partition_expect(rows, jcc, nb)
    np.testing.assert_array_equal(bnd, wbnd)
    np.testing.assert_array_equal(out, want)


def test_partition_rejects_bad_shapes():
    import pra_amd
    ctx = pra_amd.Context(nbuckets=4)
    try:
        with pytest.raises(pra_amd.PraError):
            ctx.partition_rows(0, 10, 0, 0, 0, 0)       # w = 0
        with pytest.raises(pra_amd.PraError):
            ctx.partition_rows(0, 10, 3, 4, 0, 0)       # jcc > w
        with pytest.raises(pra_amd.PraError):
            ctx.partition_rows(0, 10, 3, 1, 0, 0)       # null buffers
    finally:
        ctx.close()
