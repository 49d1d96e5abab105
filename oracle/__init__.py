"""ctypes binding of the CPU oracle (liblv_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``; the product path never imports
it.  The oracle restates the reference Louvain algorithm (modularity.cu,
aggregation.cu, main.cu) in plain C; see lv_oracle.c for the per-function
citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class OrGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("nnz", ctypes.c_uint64),
                ("row", ctypes.POINTER(ctypes.c_uint64)),
                ("col", ctypes.POINTER(ctypes.c_uint32)),
                ("w", ctypes.POINTER(ctypes.c_float))]


class OrResult(ctypes.Structure):
    _fields_ = [("q_final", ctypes.c_float), ("passes", ctypes.c_uint32),
                ("n_final", ctypes.c_uint32), ("total_iters", ctypes.c_uint32),
                ("sum_edges", ctypes.c_uint64), ("sum_edge_visits", ctypes.c_uint64),
                ("status", ctypes.c_int)]


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liblv_oracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.or_parse_mtx.argtypes = [ctypes.c_char_p, ctypes.POINTER(OrGraph)]
    L.or_parse_mtx_string.argtypes = [ctypes.c_char_p, ctypes.POINTER(OrGraph)]
    L.or_graph_free.argtypes = [ctypes.POINTER(OrGraph)]
    L.or_graph_free.restype = None
    L.or_scale_bits.argtypes = [ctypes.c_uint64, _f32p, ctypes.POINTER(ctypes.c_int)]
    L.or_r1_sum.argtypes = [_f32p, ctypes.c_uint32]
    L.or_r1_sum.restype = ctypes.c_float
    L.or_bucket_of.argtypes = [ctypes.c_uint32]
    L.or_create.argtypes = [ctypes.c_uint32, _u64p, _u32p, _f32p, ctypes.POINTER(ctypes.c_int)]
    L.or_create.restype = ctypes.c_void_p
    L.or_destroy.argtypes = [ctypes.c_void_p]
    L.or_destroy.restype = None
    L.or_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.or_pass_init.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
    L.or_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    L.or_modularity.argtypes = [ctypes.c_void_p]
    L.or_modularity.restype = ctypes.c_float
    L.or_sweep.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                           ctypes.POINTER(ctypes.c_uint64)]
    L.or_modularity_optimization.argtypes = [
        ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32)]
    L.or_aggregation.argtypes = [ctypes.c_void_p]
    L.or_run.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.POINTER(OrResult)]
    for name, rt in [("or_n", ctypes.c_uint32), ("or_ne", ctypes.c_uint64),
                     ("or_n0", ctypes.c_uint32), ("or_scale", ctypes.c_int),
                     ("or_m", ctypes.c_float)]:
        getattr(L, name).argtypes = [ctypes.c_void_p]
        getattr(L, name).restype = rt
    L.or_last_pass_q.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(ctypes.c_float)]
    L.or_last_pass_q.restype = None
    L.or_get_csr.argtypes = [ctypes.c_void_p, _u64p, _u32p, _f32p]
    L.or_get_k.argtypes = [ctypes.c_void_p, _f32p]
    L.or_get_comm.argtypes = [ctypes.c_void_p, _u32p]
    L.or_get_comm_sz.argtypes = [ctypes.c_void_p, _u32p]
    L.or_get_a.argtypes = [ctypes.c_void_p, _i64p]
    L.or_get_init_comm.argtypes = [ctypes.c_void_p, _u32p]
    for name in ["or_get_csr", "or_get_k", "or_get_comm", "or_get_comm_sz", "or_get_a",
                 "or_get_init_comm"]:
        getattr(L, name).restype = None
    L.or_q64.argtypes = [ctypes.c_uint32, _u64p, _u32p, _f32p, _u32p]
    L.or_q64.restype = ctypes.c_double
    _LIB = L
    return L


class OracleError(RuntimeError):
    pass


def _graph_to_numpy(g):
    n = g.n
    ne = g.row[n] if n >= 0 else 0
    row = np.ctypeslib.as_array(g.row, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(g.col, shape=(max(ne, 1),))[:ne].copy()
    w = np.ctypeslib.as_array(g.w, shape=(max(ne, 1),))[:ne].copy()
    return row, col, w


def parse_mtx(path=None, text=None):
    """Parse .mtx like the reference parser; returns (row u64, col u32, w f32)."""
    L = lib()
    g = OrGraph()
    if text is not None:
        rc = L.or_parse_mtx_string(text.encode(), ctypes.byref(g))
    else:
        rc = L.or_parse_mtx(str(path).encode(), ctypes.byref(g))
    if rc != 0:
        raise OracleError(f"parse failed rc={rc}")
    try:
        return _graph_to_numpy(g)
    finally:
        L.or_graph_free(ctypes.byref(g))


def r1_sum(w):
    w = np.ascontiguousarray(w, dtype=np.float32)
    return np.float32(lib().or_r1_sum(w if w.size else np.zeros(1, np.float32), w.size))


def scale_bits(w):
    w = np.ascontiguousarray(w, dtype=np.float32)
    s = ctypes.c_int(0)
    rc = lib().or_scale_bits(w.size, w if w.size else np.zeros(1, np.float32), ctypes.byref(s))
    if rc != 0:
        raise OracleError(f"scale rc={rc}")
    return s.value


def bucket_of(deg):
    return lib().or_bucket_of(int(deg))


def q64(row, col, w, assign):
    row = np.ascontiguousarray(row, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    assign = np.ascontiguousarray(assign, dtype=np.uint32)
    n = row.size - 1
    return lib().or_q64(n, row, col if col.size else np.zeros(1, np.uint32),
                        w if w.size else np.zeros(1, np.float32), assign)


class Oracle:
    """One graph's Louvain state on the CPU (mirrors the engine's handle)."""

    def __init__(self, row, col, w):
        L = lib()
        self.L = L
        row = np.ascontiguousarray(row, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        err = ctypes.c_int(0)
        n = row.size - 1
        self.h = L.or_create(n, row, col if col.size else np.zeros(1, np.uint32),
                             w if w.size else np.zeros(1, np.float32), ctypes.byref(err))
        if not self.h:
            raise OracleError(f"create failed rc={err.value}")

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(f"oracle rc={rc}")

    @property
    def n(self):
        return self.L.or_n(self.h)

    @property
    def ne(self):
        return self.L.or_ne(self.h)

    @property
    def scale(self):
        return self.L.or_scale(self.h)

    @property
    def m(self):
        return np.float32(self.L.or_m(self.h))

    def set_threads(self, nthreads):
        """Host threads for pass init, bucket steps and modularity (results do
        not depend on the count)."""
        self._chk(self.L.or_set_threads(self.h, int(nthreads)))

    def pass_init(self):
        q = ctypes.c_float(0)
        self._chk(self.L.or_pass_init(self.h, ctypes.byref(q)))
        return np.float32(q.value)

    def step(self, b):
        mv = ctypes.c_uint64(0)
        self._chk(self.L.or_step(self.h, b, ctypes.byref(mv)))
        return mv.value

    def modularity(self):
        return np.float32(self.L.or_modularity(self.h))

    def sweep(self):
        q = ctypes.c_float(0)
        mv = ctypes.c_uint64(0)
        self._chk(self.L.or_sweep(self.h, ctypes.byref(q), ctypes.byref(mv)))
        return np.float32(q.value), mv.value

    def modularity_optimization(self, min_gain):
        imp = ctypes.c_int(0)
        qp = ctypes.c_float(0)
        it = ctypes.c_uint32(0)
        self._chk(self.L.or_modularity_optimization(self.h, ctypes.c_float(min_gain),
                                                    ctypes.byref(imp), ctypes.byref(qp),
                                                    ctypes.byref(it)))
        return bool(imp.value), np.float32(qp.value), it.value

    def last_pass_q(self):
        """(Q at the last pass's init, Q after its last sweep)."""
        qb, qa = ctypes.c_float(0), ctypes.c_float(0)
        self.L.or_last_pass_q(self.h, ctypes.byref(qb), ctypes.byref(qa))
        return np.float32(qb.value), np.float32(qa.value)

    def aggregation(self):
        self._chk(self.L.or_aggregation(self.h))

    def run(self, min_gain):
        r = OrResult()
        self._chk(self.L.or_run(self.h, ctypes.c_float(min_gain), ctypes.byref(r)))
        return r

    def csr(self):
        n, ne = self.n, self.ne
        row = np.zeros(n + 1, np.uint64)
        col = np.zeros(max(ne, 1), np.uint32)
        w = np.zeros(max(ne, 1), np.float32)
        self.L.or_get_csr(self.h, row, col, w)
        return row, col[:ne], w[:ne]

    def k(self):
        a = np.zeros(max(self.n, 1), np.float32)
        self.L.or_get_k(self.h, a)
        return a[:self.n]

    def comm(self):
        a = np.zeros(max(self.n, 1), np.uint32)
        self.L.or_get_comm(self.h, a)
        return a[:self.n]

    def comm_sz(self):
        a = np.zeros(max(self.n, 1), np.uint32)
        self.L.or_get_comm_sz(self.h, a)
        return a[:self.n]

    def a_fixed(self):
        a = np.zeros(max(self.n, 1), np.int64)
        self.L.or_get_a(self.h, a)
        return a[:self.n]

    def init_comm(self):
        n0 = self.L.or_n0(self.h)
        a = np.zeros(max(n0, 1), np.uint32)
        self.L.or_get_init_comm(self.h, a)
        return a[:n0]
