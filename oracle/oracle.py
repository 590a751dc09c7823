"""ctypes binding of the CPU oracle (libpra_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this as the checker.  The product library never imports it.

The program-building calls mirror the reference's C++ surface
(relation / RAM / parallel_join / parallel_copy / parallel_acopy /
parallel_copy_filter / LIE::add_scc_dependance / LIE::execute) with integer
handles instead of pointers.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpra_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libfnv_ref.so")

DELTA, FULL, NEWT = 0, 1, 2
F_EQ, F_NEQ, F_LT, F_LE, F_GT, F_GE, F_TRUE, F_FALSE = range(8)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_intp = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_ctx_create.restype = ctypes.c_void_p
        L.ora_ctx_create.argtypes = [ctypes.c_int]
        L.ora_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.ora_set_join_threshold.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
        L.ora_set_max_iterations.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.ora_joined_tuples.restype = ctypes.c_uint64
        L.ora_joined_tuples.argtypes = [ctypes.c_void_p]
        L.ora_iterations_run.restype = ctypes.c_uint64
        L.ora_iterations_run.argtypes = [ctypes.c_void_p]
        L.ora_tuple_hash.restype = ctypes.c_uint64
        L.ora_tuple_hash.argtypes = [_u64p, ctypes.c_uint64]
        L.ora_rel_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_rel_load_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_int]
        L.ora_scc_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ora_scc_add_relation.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_scc_add_join.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, _intp, ctypes.c_int]
        for f in ("ora_scc_add_copy", "ora_scc_add_acopy"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _intp, ctypes.c_int]
        L.ora_scc_add_copy_filter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, _intp, ctypes.c_int, ctypes.c_int]
        L.ora_add_dependency.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ora_execute.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ora_scc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ora_rel_count.restype = ctypes.c_uint64
        L.ora_rel_count.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ora_rel_bucket_count.restype = ctypes.c_uint64
        L.ora_rel_bucket_count.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_rel_rows.restype = ctypes.c_uint64
        L.ora_rel_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _u64p, ctypes.c_uint64]
        L.ora_scc_history.argtypes = [ctypes.c_void_p, ctypes.c_int, _u64p, ctypes.c_int]
        L.ora_scc_bucket_history.argtypes = [ctypes.c_void_p, ctypes.c_int, _u64p, ctypes.c_uint64]
        _lib = L
    return _lib


def _ints(xs):
    arr = (ctypes.c_int * max(1, len(xs)))(*xs)
    return arr


def tuple_hash(words):
    a = np.ascontiguousarray(np.asarray(words, dtype=np.uint64))
    return int(lib().ora_tuple_hash(a.ctypes.data_as(_u64p), a.size))


def ref_tuple_hash_lib():
    """The reference's own header-only tuple_hash compiled into oracle/_ref
    (None when that build is absent, e.g. away from the reference tree)."""
    if not os.path.exists(REF_LIB_PATH):
        return None
    L = ctypes.CDLL(REF_LIB_PATH)
    L.fnv_ref_tuple_hash.restype = ctypes.c_uint64
    L.fnv_ref_tuple_hash.argtypes = [_u64p, ctypes.c_uint64]
    return L


class Oracle:
    """One LIE instance of the CPU oracle with `nbuckets` reference ranks."""

    def __init__(self, nbuckets):
        self.L = lib()
        self.c = self.L.ora_ctx_create(int(nbuckets))
        if not self.c:
            raise ValueError("bad bucket count")
        self.nb = int(nbuckets)
        self.arity = {}
        self.scc_rels = {}

    def close(self):
        if self.c:
            self.L.ora_ctx_destroy(self.c)
            self.c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_join_threshold(self, t):
        self.L.ora_set_join_threshold(self.c, int(t))

    def set_max_iterations(self, n):
        self.L.ora_set_max_iterations(self.c, int(n))

    def rel_create(self, arity, jcc, tag):
        r = self.L.ora_rel_create(self.c, arity, jcc, tag)
        if r < 0:
            raise ValueError("bad relation shape")
        self.arity[r] = arity
        return r

    def load_rows(self, rel, rows, version=FULL):
        a = np.ascontiguousarray(np.asarray(rows, dtype=np.uint64).reshape(-1, self.arity[rel] + 1))
        if self.L.ora_rel_load_rows(self.c, rel, a.ctypes.data_as(_u64p), a.shape[0], version) != 0:
            raise ValueError("load failed")

    def scc_create(self, fixpoint, sid):
        s = self.L.ora_scc_create(self.c, int(bool(fixpoint)), sid)
        self.scc_rels[s] = []
        return s

    def scc_add_relation(self, scc, rel, move_full_to_delta):
        self.L.ora_scc_add_relation(self.c, scc, rel, int(bool(move_full_to_delta)))
        self.scc_rels[scc].append(rel)

    def add_join(self, scc, out, in0, v0, in1, v1, proj):
        if self.L.ora_scc_add_join(self.c, scc, out, in0, v0, in1, v1, _ints(proj), len(proj)) != 0:
            raise ValueError("bad join")

    def add_copy(self, scc, out, src, v, reorder):
        if self.L.ora_scc_add_copy(self.c, scc, out, src, v, _ints(reorder), len(reorder)) != 0:
            raise ValueError("bad copy")

    def add_acopy(self, scc, out, src, v, reorder):
        if self.L.ora_scc_add_acopy(self.c, scc, out, src, v, _ints(reorder), len(reorder)) != 0:
            raise ValueError("bad acopy")

    def add_copy_filter(self, scc, out, src, v, reorder, op):
        if self.L.ora_scc_add_copy_filter(self.c, scc, out, src, v, _ints(reorder), len(reorder), op) != 0:
            raise ValueError("bad copy_filter")

    def add_dependency(self, before, after):
        self.L.ora_add_dependency(self.c, before, after)

    def execute(self, batch_size=1):
        self.L.ora_execute(self.c, int(batch_size))

    def count(self, rel, version=FULL):
        return int(self.L.ora_rel_count(self.c, rel, version))

    def bucket_count(self, rel, version, b):
        return int(self.L.ora_rel_bucket_count(self.c, rel, version, b))

    def rows(self, rel, version=FULL):
        n = self.count(rel, version)
        w = self.arity[rel] + 1
        out = np.zeros((max(n, 1), w), dtype=np.uint64)
        got = self.L.ora_rel_rows(self.c, rel, version, out.ctypes.data_as(_u64p), n)
        assert got == n
        return out[:n]

    def history(self, scc):
        n = self.L.ora_scc_history(self.c, scc, None, 0)
        buf = (ctypes.c_uint64 * max(1, n))()
        self.L.ora_scc_history(self.c, scc, buf, n)
        return [int(buf[i]) for i in range(n)]

    def bucket_history(self, scc):
        """per iteration, per SCC relation, per bucket delta counts"""
        nit = self.L.ora_scc_bucket_history(self.c, scc, None, 0)
        nrel = len(self.scc_rels[scc])
        tot = nit * nrel * self.nb
        buf = np.zeros(max(tot, 1), dtype=np.uint64)
        self.L.ora_scc_bucket_history(self.c, scc, buf.ctypes.data_as(_u64p), tot)
        return buf[:tot].reshape(nit, nrel, self.nb)

    def joined(self):
        return int(self.L.ora_joined_tuples(self.c))

    def iterations_run(self):
        return int(self.L.ora_iterations_run(self.c))
