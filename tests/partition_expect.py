"""CPU expectation for the stable bucket partition (numpy): bucket of a row =
FNV-1a-64 over the little-endian bytes of its first jcc words, modulo the
bucket count (the reference's tuple_hash, src/hash/hash.h:11-30, pinned by
tests/test_oracle_golden.py); rows grouped by bucket, input order kept."""
import numpy as np

FNV_OFFSET = np.uint64(14695981039346656037)
FNV_PRIME = np.uint64(1099511628211)


def fnv_rows(rows, jcc):
    rows = np.asarray(rows, dtype=np.uint64)
    h = np.full(rows.shape[0], FNV_OFFSET, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for c in range(jcc):
            v = rows[:, c]
            for b in range(8):
                h ^= (v >> np.uint64(8 * b)) & np.uint64(255)
                h *= FNV_PRIME
    return h


def partition_expect(rows, jcc, nb):
    """(rows grouped by bucket in input order, nb + 1 bucket bounds)"""
    b = fnv_rows(rows, jcc) % np.uint64(nb)
    order = np.argsort(b, kind="stable")
    bounds = np.zeros(nb + 1, dtype=np.uint64)
    bounds[1:] = np.cumsum(np.bincount(b.astype(np.int64), minlength=nb))
    return np.asarray(rows, dtype=np.uint64)[order], bounds
