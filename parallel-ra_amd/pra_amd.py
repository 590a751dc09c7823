"""Python binding of libpra_amd.so, the MI355X backend of parallel-RA's
semi-naive relational-algebra hot path.

The calls mirror the reference's C++ program surface (relation, RAM,
parallel_join / parallel_copy / parallel_acopy / parallel_copy_filter,
LIE::add_scc_dependance, LIE::execute) through the C ABI declared in
include/pra.h.  Every call goes to the HIP library; there is no CPU path, and
loading fails loudly when the library is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpra_amd.so")

DELTA, FULL, NEWT = 0, 1, 2
F_EQ, F_NEQ, F_LT, F_LE, F_GT, F_GE, F_TRUE, F_FALSE = range(8)

# every symbol include/pra.h declares
EXPORTS = (
    "pra_ctx_create", "pra_ctx_destroy", "pra_last_error", "pra_ctx_set_batch_size",
    "pra_ctx_set_join_threshold", "pra_ctx_joined_tuples", "pra_ctx_iterations",
    "pra_ctx_set_profiling", "pra_ctx_kernel_stat", "pra_ctx_synchronize", "pra_ctx_device_bytes", "pra_ctx_trim",
    "pra_ctx_set_group", "pra_ctx_group_joined", "pra_autotune_report",
    "pra_ctx_step", "pra_ctx_group_status", "pra_ctx_release_group",
    "pra_rel_create", "pra_rel_destroy", "pra_rel_load_rows", "pra_rel_load_bin", "pra_rel_count",
    "pra_rel_download", "pra_rel_write_bin", "pra_rel_checksum", "pra_rel_device_rows", "pra_scc_create", "pra_scc_destroy",
    "pra_scc_add_relation", "pra_scc_add_join", "pra_scc_add_copy", "pra_scc_add_acopy",
    "pra_scc_add_copy_filter", "pra_scc_add_dependency", "pra_scc_run", "pra_scc_begin", "pra_scc_step",
    "pra_scc_end", "pra_scc_history",
    "pra_partition_rows", "pra_partition_rows_noalloc", "pra_execute", "pra_execute_many",
    "pra_server_create", "pra_server_submit_tc", "pra_server_wait", "pra_server_rows", "pra_server_destroy",
)

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u64p = ctypes.POINTER(ctypes.c_uint64)
_intp = ctypes.POINTER(ctypes.c_int)
_lib = None


class PraError(RuntimeError):
    pass


class RequestResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("tuples", ctypes.c_uint64), ("joined", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64), ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double), ("restarts", ctypes.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    """Load libpra_amd.so (build it first with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PraError("libpra_amd.so is not built: run __graft_entry__.build() (make -C parallel-ra_amd)")
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def autotune_report():
    """the autotuner's state: [(device, knob, band, choice, [mean cost per variant...])]"""
    L = lib()
    n = _u64()
    if L.pra_autotune_report(None, 0, ctypes.byref(n)) != 0:
        raise PraError("autotune_report failed")
    buf = ctypes.create_string_buffer(int(n.value) + 1)
    L.pra_autotune_report(buf, len(buf), ctypes.byref(n))
    out = []
    for line in buf.value.decode().splitlines():
        f = line.split()
        out.append((int(f[0]), f[1], int(f[2]), int(f[3]), [float(x) for x in f[4:]]))
    return out


def load_library(path):
    """Another build of the C ABI (for example the serving loop linked with
    the CPU oracle, oracle/libpra_server_oracle.so), with the same signatures
    set on whichever of its entry points it exports."""
    return _bind(ctypes.CDLL(path))


def _bind(L):
    sig = {
        "pra_ctx_create": [ctypes.c_int, ctypes.c_int, _u64, ctypes.POINTER(_vp)],
        "pra_ctx_destroy": [_vp],
        "pra_ctx_set_batch_size": [_vp, ctypes.c_uint32],
        "pra_ctx_set_join_threshold": [_vp, ctypes.c_int64],
        "pra_ctx_joined_tuples": [_vp, _u64p],
        "pra_ctx_iterations": [_vp, _u64p],
        "pra_ctx_set_profiling": [_vp, ctypes.c_int],
        "pra_ctx_kernel_stat": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _u64p,
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
        "pra_ctx_synchronize": [_vp],
        "pra_ctx_device_bytes": [_vp, _u64p, _u64p],
        "pra_ctx_set_group": [_vp, ctypes.c_int],
        "pra_ctx_group_joined": [_vp, ctypes.c_int, _u64p],
        "pra_ctx_step": [_vp, _u64p],
        "pra_ctx_group_status": [_vp, ctypes.c_int, _intp, _u64p, _u64p],
        "pra_ctx_release_group": [_vp, ctypes.c_int],
    "pra_ctx_trim": [_vp],
    "pra_autotune_report": [ctypes.c_char_p, _u64, _u64p],
        "pra_rel_create": [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                           ctypes.POINTER(_vp)],
        "pra_rel_destroy": [_vp],
        "pra_rel_load_rows": [_vp, _u64p, _u64, ctypes.c_int],
        "pra_rel_load_bin": [_vp, ctypes.c_char_p, ctypes.c_int],
        "pra_rel_count": [_vp, ctypes.c_int, _u64p],
        "pra_rel_download": [_vp, ctypes.c_int, _u64p, _u64, _u64p],
        "pra_rel_write_bin": [_vp, ctypes.c_char_p, ctypes.c_int],
        "pra_rel_device_rows": [_vp, ctypes.c_int, ctypes.POINTER(_u64p), _u64p],
        "pra_rel_checksum": [_vp, ctypes.c_int, _u64p],
        "pra_scc_create": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
        "pra_scc_destroy": [_vp],
        "pra_scc_add_relation": [_vp, _vp, ctypes.c_int],
        "pra_scc_add_join": [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _intp, ctypes.c_int],
        "pra_scc_add_copy": [_vp, _vp, _vp, ctypes.c_int, _intp, ctypes.c_int],
        "pra_scc_add_acopy": [_vp, _vp, _vp, ctypes.c_int, _intp, ctypes.c_int],
        "pra_scc_add_copy_filter": [_vp, _vp, _vp, ctypes.c_int, _intp, ctypes.c_int, ctypes.c_int],
        "pra_scc_add_dependency": [_vp, _vp],
        "pra_scc_run": [_vp, _u64p, _u64, _u64p],
        "pra_scc_begin": [_vp, ctypes.c_int],
        "pra_scc_step": [_vp, _u64p],
        "pra_scc_end": [_vp],
        "pra_scc_history": [_vp, _u64p, _u64, _u64p],
        "pra_partition_rows": [_vp, _vp, _u64, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp],
        "pra_partition_rows_noalloc": [_vp, _vp, _u64, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp],
        "pra_execute": [_vp],
        "pra_execute_many": [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _intp],
        "pra_server_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _u64, ctypes.c_int, ctypes.POINTER(_vp)],
        "pra_server_submit_tc": [_vp, _u64p, _u64, _u64, _u64p],
        "pra_server_wait": [_vp, _u64, ctypes.POINTER(RequestResult)],
        "pra_server_rows": [_vp, _u64, _u64p, _u64, _u64p],
        "pra_server_destroy": [_vp],
    }
    for name, args in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = ctypes.c_int
    L.pra_last_error.argtypes = [_vp]
    L.pra_last_error.restype = ctypes.c_char_p
    return L


def _ints(xs):
    return (ctypes.c_int * max(1, len(xs)))(*[int(x) for x in xs])


class Context:
    """One LIE on one GPU with `nbuckets` virtual buckets (= reference ranks)."""

    def __init__(self, nbuckets=1, device=0, hbm_budget=0, batch_size=1):
        self.L = lib()
        h = _vp()
        rc = self.L.pra_ctx_create(int(device), int(nbuckets), int(hbm_budget), ctypes.byref(h))
        if rc != 0:
            raise PraError(f"pra_ctx_create failed ({rc})")
        self.c = h
        self.nb = int(nbuckets)
        self.arity = {}
        self.rels = []
        self.sccs = []
        self.scc_rels = {}
        self.cur_group = 0
        self._group_of = {}       # relation handle -> instance
        self._scc_group = {}      # scc handle -> instance
        self.set_batch_size(batch_size)

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.pra_last_error(self.c)
            raise PraError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def last_error(self):
        msg = self.L.pra_last_error(self.c)
        return msg.decode() if msg else ""

    def close(self):
        if getattr(self, "c", None):
            self.L.pra_ctx_destroy(self.c)
            self.c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_batch_size(self, bs):
        self._chk(self.L.pra_ctx_set_batch_size(self.c, int(bs)), "set_batch_size")

    def set_join_threshold(self, t):
        self._chk(self.L.pra_ctx_set_join_threshold(self.c, int(t)), "set_join_threshold")

    def set_profiling(self, on, mask=None):
        m = (-1 if on else 0) if mask is None else int(mask)
        self._chk(self.L.pra_ctx_set_profiling(self.c, m), "set_profiling")

    def kernel_names(self):
        return list(self.kernel_stats().keys())

    def kernel_stats(self):
        out = {}
        i = 0
        while True:
            name = ctypes.c_char_p()
            n = _u64()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            if self.L.pra_ctx_kernel_stat(self.c, i, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms),
                                          ctypes.byref(by)) != 0:
                break
            out[name.value.decode()] = {"launches": int(n.value), "ms": ms.value, "bytes": by.value}
            i += 1
        return out

    def synchronize(self):
        self._chk(self.L.pra_ctx_synchronize(self.c), "synchronize")

    def set_group(self, g):
        """build program instance g with the following rel_create / scc_create calls"""
        self._chk(self.L.pra_ctx_set_group(self.c, int(g)), "set_group")
        self.cur_group = int(g)

    def group_joined(self, g):
        v = _u64()
        self._chk(self.L.pra_ctx_group_joined(self.c, int(g), ctypes.byref(v)), "group_joined")
        return int(v.value)

    def step(self):
        """one lock-step iteration over the instances; returns the instances with work left"""
        v = _u64()
        self._chk(self.L.pra_ctx_step(self.c, ctypes.byref(v)), "step")
        return int(v.value)

    def group_done(self, g):
        """instance g has no SCC in progress or runnable"""
        d = ctypes.c_int()
        self._chk(self.L.pra_ctx_group_status(self.c, int(g), ctypes.byref(d), None, None), "group_status")
        return bool(d.value)

    def group_status(self, g):
        """(finished, iterations, device bytes of its relations) of instance g"""
        d, it, by = ctypes.c_int(), _u64(), _u64()
        self._chk(self.L.pra_ctx_group_status(self.c, int(g), ctypes.byref(d), ctypes.byref(it), ctypes.byref(by)),
                  "group_status")
        return bool(d.value), int(it.value), int(by.value)

    def trim(self):
        """between steps: give back the context's scratch buffers (they grow again on demand)"""
        self._chk(self.L.pra_ctx_trim(self.c), "trim")

    def release_group(self, g):
        """free instance g's relations and SCCs (their handles become invalid) and reset it"""
        self._chk(self.L.pra_ctx_release_group(self.c, int(g)), "release_group")
        self.rels = [None if self._group_of.get(i) == g else r for i, r in enumerate(self.rels)]
        self.sccs = [None if self._scc_group.get(i) == g else s for i, s in enumerate(self.sccs)]

    def device_bytes(self):
        """(bytes held now, peak bytes held) on the device by this context"""
        a, b = _u64(), _u64()
        self._chk(self.L.pra_ctx_device_bytes(self.c, ctypes.byref(a), ctypes.byref(b)), "device_bytes")
        return int(a.value), int(b.value)

    def partition_rows(self, rows, n, w, jcc, out_rows, bucket_bounds):
        """stable partition of n device rows (w words) by tuple_hash(row[0:jcc]) % nbuckets;
        rows / out_rows / bucket_bounds are device addresses (ints); ordered on the context's stream"""
        self._chk(self.L.pra_partition_rows(self.c, rows, int(n), int(w), int(jcc), out_rows, bucket_bounds),
                  "partition_rows")

    def partition_rows_noalloc(self, rows, n, w, jcc, out_rows, bucket_bounds):
        """partition_rows without device allocation (static scratch; at most 8,388,608 rows)"""
        self._chk(self.L.pra_partition_rows_noalloc(self.c, rows, int(n), int(w), int(jcc), out_rows, bucket_bounds),
                  "partition_rows_noalloc")

    def joined(self):
        v = _u64()
        self._chk(self.L.pra_ctx_joined_tuples(self.c, ctypes.byref(v)), "joined")
        return int(v.value)

    def iterations_run(self):
        v = _u64()
        self._chk(self.L.pra_ctx_iterations(self.c, ctypes.byref(v)), "iterations")
        return int(v.value)

    # relations
    def rel_create(self, arity, jcc, tag, canonical=True):
        h = _vp()
        self._chk(self.L.pra_rel_create(self.c, arity, jcc, tag, int(bool(canonical)), ctypes.byref(h)),
                  "rel_create")
        self.rels.append(h)
        r = len(self.rels) - 1
        self.arity[r] = arity
        self._group_of[r] = self.cur_group
        return r

    def load_rows(self, rel, rows, version=FULL):
        a = np.ascontiguousarray(np.asarray(rows, dtype=np.uint64).reshape(-1, self.arity[rel] + 1))
        self._chk(self.L.pra_rel_load_rows(self.rels[rel], a.ctypes.data_as(_u64p), a.shape[0], version),
                  "load_rows")

    def load_bin(self, rel, path, version=FULL):
        self._chk(self.L.pra_rel_load_bin(self.rels[rel], str(path).encode(), version), "load_bin")

    def count(self, rel, version=FULL):
        v = _u64()
        self._chk(self.L.pra_rel_count(self.rels[rel], version, ctypes.byref(v)), "count")
        return int(v.value)

    def rows(self, rel, version=FULL):
        n = self.count(rel, version)
        w = self.arity[rel] + 1
        out = np.zeros((max(n, 1), w), dtype=np.uint64)
        got = _u64()
        self._chk(self.L.pra_rel_download(self.rels[rel], version, out.ctypes.data_as(_u64p), n,
                                          ctypes.byref(got)), "download")
        assert got.value == n
        return out[:n]

    def write_bin(self, rel, path, version=FULL):
        self._chk(self.L.pra_rel_write_bin(self.rels[rel], str(path).encode(), version), "write_bin")

    def checksum(self, rel, version=FULL):
        v = _u64()
        self._chk(self.L.pra_rel_checksum(self.rels[rel], version, ctypes.byref(v)), "checksum")
        return int(v.value)

    # SCCs
    def scc_create(self, fixpoint, sid):
        h = _vp()
        self._chk(self.L.pra_scc_create(self.c, int(bool(fixpoint)), int(sid), ctypes.byref(h)), "scc_create")
        self.sccs.append(h)
        s = len(self.sccs) - 1
        self.scc_rels[s] = []
        self._scc_group[s] = self.cur_group
        return s

    def scc_add_relation(self, scc, rel, move_full_to_delta):
        self._chk(self.L.pra_scc_add_relation(self.sccs[scc], self.rels[rel], int(bool(move_full_to_delta))),
                  "scc_add_relation")
        self.scc_rels[scc].append(rel)

    def add_join(self, scc, out, in0, v0, in1, v1, proj):
        self._chk(self.L.pra_scc_add_join(self.sccs[scc], self.rels[out], self.rels[in0], v0, self.rels[in1], v1,
                                          _ints(proj), len(proj)), "add_join")

    def add_copy(self, scc, out, src, v, reorder):
        self._chk(self.L.pra_scc_add_copy(self.sccs[scc], self.rels[out], self.rels[src], v, _ints(reorder),
                                          len(reorder)), "add_copy")

    def add_acopy(self, scc, out, src, v, reorder):
        self._chk(self.L.pra_scc_add_acopy(self.sccs[scc], self.rels[out], self.rels[src], v, _ints(reorder),
                                           len(reorder)), "add_acopy")

    def add_copy_filter(self, scc, out, src, v, reorder, op):
        self._chk(self.L.pra_scc_add_copy_filter(self.sccs[scc], self.rels[out], self.rels[src], v,
                                                 _ints(reorder), len(reorder), int(op)), "add_copy_filter")

    def add_dependency(self, before, after):
        self._chk(self.L.pra_scc_add_dependency(self.sccs[before], self.sccs[after]), "add_dependency")

    def execute(self, batch_size=None):
        if batch_size is not None:
            self.set_batch_size(batch_size)
        self._chk(self.L.pra_execute(self.c), "execute")

    def history(self, scc):
        n = _u64()
        self._chk(self.L.pra_scc_history(self.sccs[scc], None, 0, ctypes.byref(n)), "history")
        buf = (_u64 * max(1, n.value))()
        self._chk(self.L.pra_scc_history(self.sccs[scc], buf, n.value, ctypes.byref(n)), "history")
        return [int(buf[i]) for i in range(n.value)]


def execute_many(contexts, concurrency=4):
    """Run independent contexts' fixed points concurrently (pra_execute_many);
    returns the per-context status codes."""
    n = len(contexts)
    arr = (_vp * max(1, n))(*[c.c for c in contexts])
    rc = (ctypes.c_int * max(1, n))()
    lib().pra_execute_many(arr, n, int(concurrency), rc)
    return [int(rc[i]) for i in range(n)]


class Server:
    """Transitive-closure request server (pra_server_*): `slots` workers,
    admission in arrival order under an HBM budget, overlapping requests."""

    def __init__(self, nbuckets=1, slots=4, hbm_budget=0, keep_rows=False, device=0, library=None):
        self.L = library if library is not None else lib()
        h = _vp()
        rc = self.L.pra_server_create(int(device), int(nbuckets), int(slots), int(hbm_budget), int(bool(keep_rows)),
                                      ctypes.byref(h))
        if rc != 0:
            raise PraError(f"pra_server_create failed ({rc})")
        self.s = h
        self._keep = {}

    def submit(self, edges, est_bytes=0):
        a = np.ascontiguousarray(np.asarray(edges, dtype=np.uint64).reshape(-1, 3))
        t = _u64()
        rc = self.L.pra_server_submit_tc(self.s, a.ctypes.data_as(_u64p), a.shape[0], int(est_bytes), ctypes.byref(t))
        if rc != 0:
            raise PraError(f"submit failed ({rc})")
        return int(t.value)

    def wait(self, ticket):
        r = RequestResult()
        rc = self.L.pra_server_wait(self.s, int(ticket), ctypes.byref(r))
        if rc != 0:
            raise PraError(f"wait failed ({rc})")
        return {f: getattr(r, f) for f, _ in RequestResult._fields_}

    def rows(self, ticket):
        n = _u64()
        if self.L.pra_server_rows(self.s, int(ticket), None, 0, ctypes.byref(n)) != 0:
            raise PraError("rows unavailable")
        out = np.zeros((max(1, n.value), 3), dtype=np.uint64)
        self.L.pra_server_rows(self.s, int(ticket), out.ctypes.data_as(_u64p), n.value, ctypes.byref(n))
        return out[:n.value]

    def close(self):
        if getattr(self, "s", None):
            self.L.pra_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
