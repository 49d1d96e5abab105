"""Python binding of the MI355X Louvain engine (liblouvain_hip.so).

Thin ctypes layer over the C ABI in include/louvain_hip.h.  The library is the
product; this module only marshals numpy arrays.  There is no CPU fallback:
if the shared library is missing, loading raises.

Because the package directory name contains dashes, import it with
``load_package()`` from this file's path or through importlib.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblouvain_hip.so")
CLI_PATH = os.path.join(HERE, "louvain")

LV_OK = 0
STATUS = {0: "ok", 1: "io", 2: "format", 3: "range", 4: "arg", 5: "nomem", 6: "state",
          7: "hip", 8: "no_device", 9: "overflow", 10: "limit"}

KERNEL_FAMILIES = ["classify", "r1", "sumsq", "move_small", "move_wave", "move_block",
                   "move_global", "apply", "aggregate", "relabel"]

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class LvError(RuntimeError):
    def __init__(self, rc, what=""):
        self.rc = rc
        super().__init__(f"{what}: status {rc} ({STATUS.get(rc, '?')})")


class HostCSR(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("nnz", ctypes.c_uint64), ("ne", ctypes.c_uint64),
                ("row", ctypes.POINTER(ctypes.c_uint32)),
                ("col", ctypes.POINTER(ctypes.c_uint32)),
                ("w", ctypes.POINTER(ctypes.c_float))]


class Result(ctypes.Structure):
    _fields_ = [("q_final", ctypes.c_float), ("passes", ctypes.c_uint32),
                ("n_final", ctypes.c_uint32), ("total_iters", ctypes.c_uint32),
                ("sum_edges", ctypes.c_uint64), ("sum_edge_visits", ctypes.c_uint64),
                ("alg_ms", ctypes.c_double), ("bytes_alg", ctypes.c_double),
                ("status", ctypes.c_int)]


class BatchGraphResult(ctypes.Structure):
    _fields_ = [("q_final", ctypes.c_float), ("passes", ctypes.c_uint32),
                ("n_final", ctypes.c_uint32), ("total_iters", ctypes.c_uint32),
                ("sum_edges", ctypes.c_uint64), ("status", ctypes.c_int)]


class RequestResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("q_final", ctypes.c_float),
                ("passes", ctypes.c_uint32), ("n_final", ctypes.c_uint32),
                ("total_iters", ctypes.c_uint32), ("sum_edges", ctypes.c_uint64),
                ("t_submit_ms", ctypes.c_double), ("t_start_ms", ctypes.c_double),
                ("t_end_ms", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("heavy_min_degree", ctypes.c_uint32), ("bucket_entries", ctypes.c_uint32),
                ("locality", ctypes.c_int32), ("locality_min_vertices", ctypes.c_uint32),
                ("wave_vertices", ctypes.c_uint32), ("block_threads", ctypes.c_uint32),
                ("class8_slots", ctypes.c_uint32), ("agg_bucket_entries", ctypes.c_uint32),
                ("sweeps_per_launch", ctypes.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelStats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 10), ("launches", ctypes.c_uint64 * 10),
                ("bytes", ctypes.c_double * 10)]


# every symbol declared in include/louvain_hip.h
EXPORTS = [
    "lv_status_string", "lv_device_count", "lv_mtx_read", "lv_host_csr_free", "lv_csr_write",
    "lv_graph_create", "lv_graph_destroy", "lv_graph_reset", "lv_graph_set_locality",
    "lv_tuning_default", "lv_graph_set_tuning", "lv_graph_get_tuning", "lv_autotune",
    "lv_graph_set_stream", "lv_modularity_optimization", "lv_pass_stats", "lv_aggregation",
    "lv_run", "lv_run_noalloc", "lv_get_assignment", "lv_graph_size", "lv_get_csr", "lv_get_state",
    "lv_get_scale", "lv_pass_init", "lv_step", "lv_sweep", "lv_modularity",
    "lv_profile_enable", "lv_profile_select", "lv_profile_get", "lv_profile_reset", "lv_gen_fem",
    "lv_gen_chung_lu", "lv_gen_rmat", "lv_gen_sbm", "lv_batch_create", "lv_batch_run",
    "lv_batch_get_assignment", "lv_graph_footprint", "lv_server_create", "lv_server_submit",
    "lv_server_poll", "lv_server_get_assignment", "lv_server_drain", "lv_server_destroy",
]

_LIB = None


def build(jobs=8):
    """Compile liblouvain_hip.so and the CLI for gfx950 (always rebuilds)."""
    subprocess.run(["make", "-B", f"-j{jobs}", "-C", HERE], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    c = ctypes
    L.lv_status_string.argtypes = [c.c_int]
    L.lv_status_string.restype = c.c_char_p
    L.lv_device_count.argtypes = [c.POINTER(c.c_int)]
    L.lv_mtx_read.argtypes = [c.c_char_p, c.POINTER(HostCSR)]
    L.lv_host_csr_free.argtypes = [c.POINTER(HostCSR)]
    L.lv_host_csr_free.restype = None
    L.lv_csr_write.argtypes = [c.c_char_p, c.POINTER(HostCSR)]
    L.lv_graph_create.argtypes = [c.c_int, c.c_uint32, _u32p, _u32p, _f32p,
                                  c.POINTER(c.c_void_p)]
    L.lv_graph_destroy.argtypes = [c.c_void_p]
    L.lv_graph_reset.argtypes = [c.c_void_p]
    L.lv_graph_set_locality.argtypes = [c.c_void_p, c.c_int]
    L.lv_tuning_default.argtypes = [c.POINTER(Tuning)]
    L.lv_graph_set_tuning.argtypes = [c.c_void_p, c.POINTER(Tuning)]
    L.lv_graph_get_tuning.argtypes = [c.c_void_p, c.POINTER(Tuning)]
    L.lv_autotune.argtypes = [c.c_void_p, c.c_float, c.POINTER(Tuning), c.c_uint32,
                              c.POINTER(Tuning), c.POINTER(c.c_double)]
    L.lv_graph_set_stream.argtypes = [c.c_void_p, c.c_void_p]
    L.lv_modularity_optimization.argtypes = [c.c_void_p, c.c_float, c.POINTER(c.c_int),
                                             c.POINTER(c.c_float), c.POINTER(c.c_float)]
    L.lv_pass_stats.argtypes = [c.c_void_p, c.POINTER(c.c_uint32), c.POINTER(c.c_float),
                                c.POINTER(c.c_float)]
    L.lv_aggregation.argtypes = [c.c_void_p]
    L.lv_run.argtypes = [c.c_void_p, c.c_float, c.POINTER(Result)]
    L.lv_run_noalloc.argtypes = [c.c_int, c.c_uint32, _u32p, _u32p, _f32p, c.c_float,
                                 c.c_void_p, _u32p, c.POINTER(Result)]
    L.lv_get_assignment.argtypes = [c.c_void_p, _u32p]
    L.lv_graph_size.argtypes = [c.c_void_p, c.POINTER(c.c_uint32), c.POINTER(c.c_uint64),
                                c.POINTER(c.c_uint32)]
    L.lv_get_csr.argtypes = [c.c_void_p, _u32p, _u32p, _f32p]
    L.lv_get_state.argtypes = [c.c_void_p, _f32p, _u32p, _u32p, _i64p]
    L.lv_get_scale.argtypes = [c.c_void_p, c.POINTER(c.c_int), c.POINTER(c.c_float)]
    L.lv_pass_init.argtypes = [c.c_void_p, c.POINTER(c.c_float)]
    L.lv_step.argtypes = [c.c_void_p, c.c_int, c.POINTER(c.c_uint64)]
    L.lv_sweep.argtypes = [c.c_void_p, c.POINTER(c.c_float), c.POINTER(c.c_uint64)]
    L.lv_modularity.argtypes = [c.c_void_p, c.POINTER(c.c_float)]
    L.lv_profile_enable.argtypes = [c.c_void_p, c.c_int]
    L.lv_profile_select.argtypes = [c.c_void_p, c.c_uint32]
    L.lv_profile_get.argtypes = [c.c_void_p, c.POINTER(KernelStats)]
    L.lv_profile_reset.argtypes = [c.c_void_p]
    L.lv_gen_fem.argtypes = [c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint64, c.POINTER(HostCSR)]
    L.lv_gen_chung_lu.argtypes = [c.c_uint32, c.c_double, c.c_double, c.c_uint32, c.c_uint32,
                                  c.c_double, c.c_uint64, c.POINTER(HostCSR)]
    L.lv_gen_rmat.argtypes = [c.c_uint32, c.c_uint32, c.c_double, c.c_double, c.c_double,
                              c.c_float, c.c_float, c.c_uint64, c.POINTER(HostCSR)]
    L.lv_gen_sbm.argtypes = [c.c_uint32, c.c_uint32, c.c_double, c.c_double, c.c_uint64,
                             c.POINTER(HostCSR)]
    _u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
    L.lv_batch_create.argtypes = [c.c_int, c.c_uint32, _u64p, _u32p, _u32p, _f32p,
                                  c.POINTER(c.c_void_p)]
    L.lv_batch_run.argtypes = [c.c_void_p, c.c_float, c.POINTER(BatchGraphResult),
                               c.POINTER(Result)]
    L.lv_batch_get_assignment.argtypes = [c.c_void_p, _u32p]
    L.lv_graph_footprint.argtypes = [c.c_uint32, c.c_uint64]
    L.lv_graph_footprint.restype = c.c_uint64
    L.lv_server_create.argtypes = [c.c_int, c.c_uint32, c.c_uint64, c.c_float, c.c_int,
                                   c.POINTER(c.c_void_p)]
    L.lv_server_submit.argtypes = [c.c_void_p, c.c_uint64, c.c_uint32, _u32p, _u32p, _f32p]
    L.lv_server_poll.argtypes = [c.c_void_p, c.POINTER(RequestResult), c.c_uint32,
                                 c.POINTER(c.c_uint32)]
    L.lv_server_get_assignment.argtypes = [c.c_void_p, c.c_uint64, _u32p, c.c_uint32]
    L.lv_server_drain.argtypes = [c.c_void_p]
    L.lv_server_destroy.argtypes = [c.c_void_p]
    _LIB = L
    return L


def _chk(rc, what):
    if rc != LV_OK:
        raise LvError(rc, what)


def device_count():
    n = ctypes.c_int(0)
    rc = lib().lv_device_count(ctypes.byref(n))
    return n.value if rc == LV_OK else 0


def _take_csr(h):
    n, ne = h.n, h.ne
    row = np.ctypeslib.as_array(h.row, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(h.col, shape=(max(ne, 1),))[:ne].copy()
    w = np.ctypeslib.as_array(h.w, shape=(max(ne, 1),))[:ne].copy()
    lib().lv_host_csr_free(ctypes.byref(h))
    return row, col, w


def read_mtx(path):
    """lv_mtx_read: (row u32[n+1], col u32[ne], w f32[ne]) like utils.cu parseGraph."""
    h = HostCSR()
    _chk(lib().lv_mtx_read(str(path).encode(), ctypes.byref(h)), "lv_mtx_read")
    return _take_csr(h)


def write_csr(path, row, col, w, nnz=0):
    """lv_csr_write: binary CSR cache that read_mtx (lv_mtx_read) loads."""
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    h = HostCSR()
    h.n, h.nnz, h.ne = row.size - 1, nnz, col.size
    h.row = row.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    h.col = col.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    h.w = w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    _chk(lib().lv_csr_write(str(path).encode(), ctypes.byref(h)), "lv_csr_write")


def gen_fem(nodes=814, band=25, couple=9, seed=1):
    h = HostCSR()
    _chk(lib().lv_gen_fem(nodes, band, couple, seed, ctypes.byref(h)), "lv_gen_fem")
    return _take_csr(h)


def gen_chung_lu(n=262111, mean_deg=9.4, exponent=2.5, max_deg=420, groups=2000, p_ratio=20.0,
                 seed=2):
    h = HostCSR()
    _chk(lib().lv_gen_chung_lu(n, mean_deg, exponent, max_deg, groups, p_ratio, seed,
                               ctypes.byref(h)), "lv_gen_chung_lu")
    return _take_csr(h)


def gen_rmat(scale=24, edge_factor=16, a=0.57, b=0.19, c=0.19, wlo=0.5, whi=1.5, seed=3):
    h = HostCSR()
    _chk(lib().lv_gen_rmat(scale, edge_factor, a, b, c, wlo, whi, seed, ctypes.byref(h)),
         "lv_gen_rmat")
    return _take_csr(h)


def gen_sbm(n, blocks, mean_deg, p_ratio=10.0, seed=4):
    h = HostCSR()
    _chk(lib().lv_gen_sbm(n, blocks, mean_deg, p_ratio, seed, ctypes.byref(h)), "lv_gen_sbm")
    return _take_csr(h)


def run_noalloc(row, col, w, min_gain=1e-6, device=0, stream=None):
    """lv_run_noalloc: one whole run with every device buffer carved from the
    library's fixed scratch array.  Returns (Result, assignment)."""
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    n = row.size - 1
    a = np.zeros(max(n, 1), np.uint32)
    r = Result()
    _chk(lib().lv_run_noalloc(device, n, row, col if col.size else np.zeros(1, np.uint32),
                              w if w.size else np.zeros(1, np.float32), ctypes.c_float(min_gain),
                              ctypes.c_void_p(stream or None), a, ctypes.byref(r)),
         "lv_run_noalloc")
    return r, a[:n]


class _Tunable:
    """lv_graph_get/set_tuning and lv_autotune for graph and batch handles."""

    def get_tuning(self):
        t = Tuning()
        _chk(self.L.lv_graph_get_tuning(self.h, ctypes.byref(t)), "lv_graph_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning(**{**self.get_tuning(), **kw})
        _chk(self.L.lv_graph_set_tuning(self.h, ctypes.byref(t)), "lv_graph_set_tuning")

    def autotune(self, min_gain=1e-6, candidates=None):
        """lv_autotune over `candidates` (dicts) or the built-in grid; returns
        (best tuning dict, best ms); the graph keeps the best and is reset."""
        cands = None
        n = 0
        if candidates:
            base = self.get_tuning()
            cands = (Tuning * len(candidates))(*[Tuning(**{**base, **c}) for c in candidates])
            n = len(candidates)
        best, ms = Tuning(), ctypes.c_double()
        _chk(self.L.lv_autotune(self.h, ctypes.c_float(min_gain), cands, n, ctypes.byref(best),
                                ctypes.byref(ms)), "lv_autotune")
        return best.as_dict(), ms.value


class Graph(_Tunable):
    """A device-resident graph (lv_graph*) with the reference operator API."""

    def __init__(self, row, col, w, device=0):
        L = lib()
        self.L = L
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        h = ctypes.c_void_p()
        rc = L.lv_graph_create(device, row.size - 1, row,
                               col if col.size else np.zeros(1, np.uint32),
                               w if w.size else np.zeros(1, np.float32), ctypes.byref(h))
        _chk(rc, "lv_graph_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.lv_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reset(self):
        _chk(self.L.lv_graph_reset(self.h), "lv_graph_reset")

    def set_locality(self, mode):
        """lv_graph_set_locality: -1 automatic, 0 never, 1 every pass."""
        _chk(self.L.lv_graph_set_locality(self.h, int(mode)), "lv_graph_set_locality")


    def size(self):
        n, ne, n0 = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32()
        _chk(self.L.lv_graph_size(self.h, ctypes.byref(n), ctypes.byref(ne), ctypes.byref(n0)),
             "lv_graph_size")
        return n.value, ne.value, n0.value

    def pass_init(self):
        q = ctypes.c_float()
        _chk(self.L.lv_pass_init(self.h, ctypes.byref(q)), "lv_pass_init")
        return np.float32(q.value)

    def step(self, b):
        mv = ctypes.c_uint64()
        _chk(self.L.lv_step(self.h, b, ctypes.byref(mv)), "lv_step")
        return mv.value

    def sweep(self):
        q, mv = ctypes.c_float(), ctypes.c_uint64()
        _chk(self.L.lv_sweep(self.h, ctypes.byref(q), ctypes.byref(mv)), "lv_sweep")
        return np.float32(q.value), mv.value

    def modularity(self):
        q = ctypes.c_float()
        _chk(self.L.lv_modularity(self.h, ctypes.byref(q)), "lv_modularity")
        return np.float32(q.value)

    def modularity_optimization(self, min_gain):
        """One pass of local moves: (improved, q_before, q_after, sweeps)."""
        imp, qb, qa = ctypes.c_int(), ctypes.c_float(), ctypes.c_float()
        _chk(self.L.lv_modularity_optimization(self.h, ctypes.c_float(min_gain),
                                               ctypes.byref(imp), ctypes.byref(qb),
                                               ctypes.byref(qa)),
             "lv_modularity_optimization")
        return bool(imp.value), np.float32(qb.value), np.float32(qa.value), self.pass_stats()[0]

    def pass_stats(self):
        """(sweeps, q_before, q_after) of the last pass."""
        it, qb, qa = ctypes.c_uint32(), ctypes.c_float(), ctypes.c_float()
        _chk(self.L.lv_pass_stats(self.h, ctypes.byref(it), ctypes.byref(qb), ctypes.byref(qa)),
             "lv_pass_stats")
        return it.value, np.float32(qb.value), np.float32(qa.value)

    def set_stream(self, stream):
        """Run later calls on a caller's HIP stream (an integer handle such as
        torch.cuda.Stream().cuda_stream); 0 or None returns to the graph's own."""
        _chk(self.L.lv_graph_set_stream(self.h, ctypes.c_void_p(stream or None)),
             "lv_graph_set_stream")

    def aggregation(self):
        _chk(self.L.lv_aggregation(self.h), "lv_aggregation")

    def run(self, min_gain):
        r = Result()
        _chk(self.L.lv_run(self.h, ctypes.c_float(min_gain), ctypes.byref(r)), "lv_run")
        return r

    def assignment(self):
        n0 = self.size()[2]
        a = np.zeros(max(n0, 1), np.uint32)
        _chk(self.L.lv_get_assignment(self.h, a), "lv_get_assignment")
        return a[:n0]

    def csr(self):
        n, ne, _ = self.size()
        row = np.zeros(n + 1, np.uint32)
        col = np.zeros(max(ne, 1), np.uint32)
        w = np.zeros(max(ne, 1), np.float32)
        _chk(self.L.lv_get_csr(self.h, row, col, w), "lv_get_csr")
        return row, col[:ne], w[:ne]

    def state(self):
        n = self.size()[0]
        k = np.zeros(max(n, 1), np.float32)
        comm = np.zeros(max(n, 1), np.uint32)
        sz = np.zeros(max(n, 1), np.uint32)
        a = np.zeros(max(n, 1), np.int64)
        _chk(self.L.lv_get_state(self.h, k, comm, sz, a), "lv_get_state")
        return k[:n], comm[:n], sz[:n], a[:n]

    def scale(self):
        s, m = ctypes.c_int(), ctypes.c_float()
        _chk(self.L.lv_get_scale(self.h, ctypes.byref(s), ctypes.byref(m)), "lv_get_scale")
        return s.value, np.float32(m.value)

    def profile(self, on=True):
        _chk(self.L.lv_profile_enable(self.h, 1 if on else 0), "lv_profile_enable")

    def profile_select(self, families):
        """Profile only the named kernel families (None: all)."""
        mask = 0xFFFFFFFF if families is None else sum(1 << KERNEL_FAMILIES.index(f)
                                                        for f in families)
        _chk(self.L.lv_profile_select(self.h, mask), "lv_profile_select")

    def profile_reset(self):
        _chk(self.L.lv_profile_reset(self.h), "lv_profile_reset")

    def profile_get(self):
        st = KernelStats()
        _chk(self.L.lv_profile_get(self.h, ctypes.byref(st)), "lv_profile_get")
        return {f: {"ms": st.ms[i], "launches": st.launches[i], "bytes": st.bytes[i]}
                for i, f in enumerate(KERNEL_FAMILIES)}


def concat_graphs(graphs):
    """Concatenate [(row, col, w), ...] into one segmented CSR with global ids:
    returns (graph_off u64[G+1], row u32[n+1], col u32[ne], w f32[ne])."""
    sizes = [int(r.size) - 1 for r, _, _ in graphs]
    graph_off = np.zeros(len(graphs) + 1, np.uint64)
    graph_off[1:] = np.cumsum(sizes)
    rows, cols, ws = [np.zeros(1, np.uint64)], [], []
    ebase = 0
    for (r, c, w), vo in zip(graphs, graph_off[:-1]):
        r = r.astype(np.uint64)
        rows.append(r[1:] + ebase)
        cols.append(c.astype(np.uint64) + vo)
        ws.append(w.astype(np.float32))
        ebase += int(r[-1])
    row = np.concatenate(rows)
    if row[-1] >= 2 ** 32:
        raise ValueError("batch exceeds 2^32 entries")
    col = np.concatenate(cols) if cols else np.zeros(0, np.uint64)
    return (graph_off, row.astype(np.uint32), col.astype(np.uint32),
            np.concatenate(ws) if ws else np.zeros(0, np.float32))


class Batch(_Tunable):
    """Many graphs processed together (lv_batch_*); results per graph equal
    single-graph runs bit for bit."""

    def __init__(self, graphs, device=0):
        L = lib()
        self.L = L
        self.graph_off, row, col, w = concat_graphs(graphs)
        self.G = len(graphs)
        h = ctypes.c_void_p()
        _chk(L.lv_batch_create(device, self.G, self.graph_off, row,
                               col if col.size else np.zeros(1, np.uint32),
                               w if w.size else np.zeros(1, np.float32), ctypes.byref(h)),
             "lv_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.lv_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reset(self):
        _chk(self.L.lv_graph_reset(self.h), "lv_graph_reset")

    def run(self, min_gain):
        per = (BatchGraphResult * self.G)()
        tot = Result()
        _chk(self.L.lv_batch_run(self.h, ctypes.c_float(min_gain), per, ctypes.byref(tot)),
             "lv_batch_run")
        return list(per), tot

    def assignments(self):
        n0 = int(self.graph_off[-1])
        a = np.zeros(max(n0, 1), np.uint32)
        _chk(self.L.lv_batch_get_assignment(self.h, a), "lv_batch_get_assignment")
        return [a[int(self.graph_off[i]):int(self.graph_off[i + 1])] for i in range(self.G)]

    def profile(self, on=True):
        _chk(self.L.lv_profile_enable(self.h, 1 if on else 0), "lv_profile_enable")

    def profile_get(self):
        st = KernelStats()
        _chk(self.L.lv_profile_get(self.h, ctypes.byref(st)), "lv_profile_get")
        return {f: {"ms": st.ms[i], "launches": st.launches[i], "bytes": st.bytes[i]}
                for i, f in enumerate(KERNEL_FAMILIES)}


class Server:
    """Request loop with continuous admission (lv_server_*)."""

    def __init__(self, slots=8, hbm_budget=200 << 30, min_gain=1e-6, keep_assignments=False,
                 device=0):
        L = lib()
        self.L = L
        h = ctypes.c_void_p()
        _chk(L.lv_server_create(device, slots, hbm_budget, ctypes.c_float(min_gain),
                                1 if keep_assignments else 0, ctypes.byref(h)),
             "lv_server_create")
        self.h = h

    def submit(self, rid, row, col, w):
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        _chk(self.L.lv_server_submit(self.h, rid, row.size - 1, row,
                                     col if col.size else np.zeros(1, np.uint32),
                                     w if w.size else np.zeros(1, np.float32)),
             "lv_server_submit")

    def poll(self, max_results=1024):
        buf = (RequestResult * max_results)()
        cnt = ctypes.c_uint32()
        _chk(self.L.lv_server_poll(self.h, buf, max_results, ctypes.byref(cnt)), "lv_server_poll")
        return [buf[i] for i in range(cnt.value)]

    def assignment(self, rid, n):
        a = np.zeros(max(n, 1), np.uint32)
        _chk(self.L.lv_server_get_assignment(self.h, rid, a, n), "lv_server_get_assignment")
        return a[:n]

    def drain(self):
        _chk(self.L.lv_server_drain(self.h), "lv_server_drain")

    def close(self):
        if getattr(self, "h", None):
            self.L.lv_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
