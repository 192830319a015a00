"""MI355X signed triangle counter: Python binding of the C ABI in include/stc.h.

The native library ``libstc.so`` (built in-tree by ``build.py``) does all the
work: on-device CSR build and degree orientation, binning, and the HIP merge /
LDS-hash / hub-bitmap intersection kernels.  This module only marshals
arguments through ctypes; there is no Python or CPU fallback, so a missing
library raises at import of the first call.

Reference interface mirrored (Signed_TC/GPU Implementations/GlobalMEM/
signed_tc.hxx:17-24,527-542): ``signed_tc.run(G, param, result)`` writes
``result.total_balanced`` / ``result.total_unbalanced`` and returns the
counting time in ms.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstc.so")
CLI_PATH = os.path.join(_HERE, "stc")
MODEL_DIR = os.path.join(_HERE, "models")

STC_OK = 0
STC_E_ARG = -1
STC_E_TOO_LARGE = -2
STC_E_OOM = -3
STC_E_HIP = -4
STC_E_MODEL = -5
STC_E_IO = -6
STC_E_NODEV = -7

DISPATCH = {"learned": 0, "fixed": 1, "merge": 2, "hash": 3, "bitmap": 4, "pull": 5}

KERNELS = ["merge_lane", "merge_wave", "hash_wave", "hash", "bitmap_wave", "bitmap",
           "pull_wave", "pull"]

# Names every build of libstc.so must export (the declarations of
# include/stc.h and include/stc_gen.h).
EXPORTED = [
    "stc_create", "stc_destroy", "stc_last_error", "stc_model_status",
    "stc_count", "stc_count_device", "stc_last_classes", "stc_prepare",
    "stc_count_prepared", "stc_count_prepared_static", "stc_count_batch", "stc_prepare_batch",
    "stc_count_batch_prepared", "stc_count_per_vertex", "stc_get_stats", "stc_run",
    "stc_load_graph", "stc_parse_graph", "stc_write_graph", "stc_graph_free",
    "stc_gen_er", "stc_gen_grid", "stc_gen_rmat", "stc_gen_hub", "stc_gen_batch",
    "stc_batch_free", "stc_gen_workload", "stc_gen_workload_batch",
    "stc_bucket_histogram", "stc_count_with_table", "stc_bucket_features",
    "stc_train_dispatcher", "stc_load_models", "stc_set_tune", "stc_last_table",
    "stc_get_csr",
]

TABLE = {"skip": 0, "merge": 1, "hash": 2, "bitmap": 3, "pull": 4}


class StcError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__("stc error %d%s" % (code, (": " + msg) if msg else ""))
        self.code = code


class _HostGraph(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64),
                ("u", C.POINTER(C.c_uint32)), ("v", C.POINTER(C.c_uint32)),
                ("sign", C.POINTER(C.c_int8))]


class _HostBatch(C.Structure):
    _fields_ = [("k", C.c_uint32), ("total_m", C.c_uint64),
                ("edge_off", C.POINTER(C.c_uint64)), ("n", C.POINTER(C.c_uint32)),
                ("kind", C.POINTER(C.c_uint8)), ("p_neg", C.POINTER(C.c_double)),
                ("avg_deg", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_uint32)),
                ("v", C.POINTER(C.c_uint32)), ("sign", C.POINTER(C.c_int8))]


class Edges(C.Structure):
    """stc_edges: n, m and the u / v / sign pointers (host or device)."""
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64),
                ("u", C.c_void_p), ("v", C.c_void_p), ("sign", C.c_void_p)]


class Counts(C.Structure):
    _fields_ = [("balanced", C.c_uint64), ("unbalanced", C.c_uint64)]


class Classes(C.Structure):
    _fields_ = [("t0", C.c_uint64), ("t1", C.c_uint64), ("t2", C.c_uint64),
                ("t3", C.c_uint64)]


class Opts(C.Structure):
    _fields_ = [("dispatch", C.c_int), ("repeat", C.c_int), ("classes", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m_in", C.c_uint64), ("m_oriented", C.c_uint64),
                ("max_out_degree", C.c_uint32), ("b_alg", C.c_double),
                ("ms_upload", C.c_float), ("ms_build", C.c_float),
                ("ms_count", C.c_float), ("ms_count_min", C.c_float),
                ("ms_bin", C.c_float), ("edges_merge", C.c_uint64),
                ("edges_hash", C.c_uint64), ("edges_bitmap", C.c_uint64), ("edges_pull", C.c_uint64),
                ("edges_skipped", C.c_uint64), ("dispatch_used", C.c_int),
                ("ms_kernel", C.c_float * 8), ("bytes_kernel", C.c_double * 8),
                ("alg_bytes_kernel", C.c_double * 8)]

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = list(v) if k in ("ms_kernel", "bytes_kernel", "alg_bytes_kernel") else v
        return d


_lib_handle = None


def lib():
    """Load libstc.so (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise ImportError("libstc.so not built: run signed_triangle_counting_amd/build.py "
                          "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.stc_create.argtypes = [C.c_int, C.c_char_p, P(C.c_void_p)]
    L.stc_destroy.argtypes = [C.c_void_p]
    L.stc_destroy.restype = None
    L.stc_last_error.argtypes = [C.c_void_p]
    L.stc_last_error.restype = C.c_char_p
    L.stc_model_status.argtypes = [C.c_void_p]
    L.stc_count.argtypes = [C.c_void_p, P(Edges), P(Opts), P(Counts)]
    L.stc_count_device.argtypes = [C.c_void_p, P(Edges), P(Opts), P(Counts)]
    L.stc_prepare.argtypes = [C.c_void_p, P(Edges), C.c_int]
    L.stc_count_prepared.argtypes = [C.c_void_p, P(Opts), P(Counts)]
    L.stc_count_prepared_static.argtypes = [C.c_void_p, P(Opts), P(Counts)]
    L.stc_count_batch.argtypes = [C.c_void_p, P(Edges), C.c_uint32, P(Opts), P(Counts)]
    L.stc_prepare_batch.argtypes = [C.c_void_p, P(Edges), C.c_uint32]
    L.stc_count_per_vertex.argtypes = [C.c_void_p, P(Opts), C.c_void_p, P(Counts)]
    L.stc_count_batch_prepared.argtypes = [C.c_void_p, P(Opts), P(Counts)]
    L.stc_last_classes.argtypes = [C.c_void_p, P(Classes)]
    L.stc_get_stats.argtypes = [C.c_void_p, P(Stats)]
    L.stc_run.argtypes = [C.c_void_p, P(Edges), P(C.c_uint64), P(C.c_uint64)]
    L.stc_run.restype = C.c_float
    L.stc_load_graph.argtypes = [C.c_char_p, C.c_int, P(P(_HostGraph))]
    L.stc_parse_graph.argtypes = [C.c_char_p, C.c_uint64, C.c_int, P(P(_HostGraph))]
    L.stc_write_graph.argtypes = [C.c_char_p, P(_HostGraph)]
    L.stc_graph_free.argtypes = [P(_HostGraph)]
    L.stc_graph_free.restype = None
    L.stc_batch_free.argtypes = [P(_HostBatch)]
    L.stc_batch_free.restype = None
    L.stc_gen_er.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_uint64, P(P(_HostGraph))]
    L.stc_gen_grid.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_uint64,
                               P(P(_HostGraph))]
    L.stc_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_uint64, C.c_int, P(P(_HostGraph))]
    L.stc_gen_hub.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                              C.c_uint32, C.c_double, C.c_uint64, C.c_int, P(P(_HostGraph))]
    L.stc_gen_batch.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_uint64, C.c_int, P(P(_HostBatch))]
    L.stc_gen_workload.argtypes = [C.c_int, C.c_int, P(P(_HostGraph))]
    L.stc_gen_workload_batch.argtypes = [C.c_uint32, C.c_int, P(P(_HostBatch))]
    L.stc_bucket_histogram.argtypes = [C.c_void_p, C.c_void_p]
    L.stc_count_with_table.argtypes = [C.c_void_p, C.c_void_p, C.c_int, P(Counts)]
    L.stc_bucket_features.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                      C.c_void_p]
    L.stc_bucket_features.restype = None
    L.stc_train_dispatcher.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                       C.c_float, C.c_uint64, C.c_char_p]
    L.stc_load_models.argtypes = [C.c_void_p, C.c_char_p]
    L.stc_set_tune.argtypes = [C.c_void_p, C.c_char_p]
    L.stc_last_table.argtypes = [C.c_void_p, C.c_void_p]
    L.stc_get_csr.argtypes = [C.c_void_p] + [C.c_void_p] * 7
    _lib_handle = L
    return L


def _check(rc, ctx=None):
    if rc != STC_OK:
        msg = ""
        if ctx:
            m = lib().stc_last_error(ctx)
            msg = m.decode() if m else ""
        raise StcError(rc, msg)


class Graph:
    """A host edge list in library-owned memory (or wrapping numpy arrays)."""

    def __init__(self, ptr=None, u=None, v=None, sign=None, n=None):
        self._ptr = ptr
        if ptr is not None:
            g = ptr.contents
            self.n, self.m = int(g.n), int(g.m)
            if self.m:
                self.u = np.ctypeslib.as_array(g.u, (self.m,))
                self.v = np.ctypeslib.as_array(g.v, (self.m,))
                self.sign = np.ctypeslib.as_array(g.sign, (self.m,))
            else:
                self.u = np.zeros(0, np.uint32)
                self.v = np.zeros(0, np.uint32)
                self.sign = np.zeros(0, np.int8)
        else:
            self.u = np.ascontiguousarray(u, dtype=np.uint32)
            self.v = np.ascontiguousarray(v, dtype=np.uint32)
            self.sign = np.ascontiguousarray(sign, dtype=np.int8)
            self.m = int(self.u.shape[0])
            if n is None:
                n = int(max(self.u.max(initial=0), self.v.max(initial=0))) + 1 if self.m else 0
            self.n = int(n)

    def edges(self):
        return Edges(self.n, self.m, self.u.ctypes.data if self.m else None,
                     self.v.ctypes.data if self.m else None,
                     self.sign.ctypes.data if self.m else None)

    def write(self, path):
        if self._ptr is not None:
            _check(lib().stc_write_graph(path.encode(), self._ptr))
        else:
            with open(path, "w") as f:
                for a, b, s in zip(self.u.tolist(), self.v.tolist(), self.sign.tolist()):
                    f.write("%d %d %s\n" % (a, b, "-1" if s < 0 else "1"))

    def __del__(self):
        if getattr(self, "_ptr", None) is not None and _lib_handle is not None:
            _lib_handle.stc_graph_free(self._ptr)
            self._ptr = None


class Batch:
    """k graphs in one pool (stc_host_batch)."""

    def __init__(self, ptr):
        self._ptr = ptr
        b = ptr.contents
        self.k = int(b.k)
        self.total_m = int(b.total_m)
        self.edge_off = np.ctypeslib.as_array(b.edge_off, (self.k + 1,))
        self.n = np.ctypeslib.as_array(b.n, (self.k,))
        self.kind = np.ctypeslib.as_array(b.kind, (self.k,))
        self.p_neg = np.ctypeslib.as_array(b.p_neg, (self.k,))
        self.avg_deg = np.ctypeslib.as_array(b.avg_deg, (self.k,))
        tm = max(self.total_m, 1)
        self.u = np.ctypeslib.as_array(b.u, (tm,))[: self.total_m]
        self.v = np.ctypeslib.as_array(b.v, (tm,))[: self.total_m]
        self.sign = np.ctypeslib.as_array(b.sign, (tm,))[: self.total_m]

    def graph_edges(self, g):
        lo, hi = int(self.edge_off[g]), int(self.edge_off[g + 1])
        m = hi - lo
        base_u = self.u.ctypes.data + 4 * lo
        base_v = self.v.ctypes.data + 4 * lo
        base_s = self.sign.ctypes.data + lo
        return Edges(int(self.n[g]), m, base_u, base_v, base_s)

    def graph(self, g):
        lo, hi = int(self.edge_off[g]), int(self.edge_off[g + 1])
        return Graph(u=self.u[lo:hi], v=self.v[lo:hi], sign=self.sign[lo:hi], n=int(self.n[g]))

    def __del__(self):
        if getattr(self, "_ptr", None) is not None and _lib_handle is not None:
            _lib_handle.stc_batch_free(self._ptr)
            self._ptr = None


def _graph_out(fn, *args):
    p = C.POINTER(_HostGraph)()
    _check(fn(*args, C.byref(p)))
    return Graph(ptr=p)


def load_graph(path, threads=0):
    """Parse a graph file in the reference text format (graph_io.h:72-212)."""
    return _graph_out(lib().stc_load_graph, path.encode(), threads)


def parse_graph(text, threads=0):
    data = text.encode() if isinstance(text, str) else bytes(text)
    return _graph_out(lib().stc_parse_graph, data, len(data), threads)


def gen_er(n, m, p_neg, seed):
    return _graph_out(lib().stc_gen_er, n, m, p_neg, seed)


def gen_grid(rows, cols, p_diag, p_neg, seed):
    return _graph_out(lib().stc_gen_grid, rows, cols, p_diag, p_neg, seed)


def gen_rmat(scale, edge_factor, a, b, c, p_neg, seed, threads=0):
    return _graph_out(lib().stc_gen_rmat, scale, edge_factor, a, b, c, p_neg, seed, threads)


def gen_hub(n, target_m, hubs, zipf_s, hub_share, planted, p_neg, seed, threads=0):
    return _graph_out(lib().stc_gen_hub, n, target_m, hubs, zipf_s, hub_share, planted,
                      p_neg, seed, threads)


def gen_workload(wid, threads=0):
    """Benchmark workloads 1..4 (ER 10k/100k, grid 960^2, RMAT-22, hub 18M)."""
    return _graph_out(lib().stc_gen_workload, wid, threads)


def gen_batch(k, n_lo=64, n_hi=65536, d_lo=4.0, d_hi=64.0, p_lo=0.05, p_hi=0.5,
              seed=1, threads=0):
    p = C.POINTER(_HostBatch)()
    _check(lib().stc_gen_batch(k, n_lo, n_hi, d_lo, d_hi, p_lo, p_hi, seed, threads,
                               C.byref(p)))
    return Batch(p)


def gen_workload_batch(k=4096, threads=0):
    p = C.POINTER(_HostBatch)()
    _check(lib().stc_gen_workload_batch(k, threads, C.byref(p)))
    return Batch(p)


def _as_graph(g, u=None, v=None, sign=None, n=None):
    if isinstance(g, Graph):
        return g
    if g is None:
        return Graph(u=u, v=v, sign=sign, n=n)
    raise TypeError("expected a Graph")


class Counter:
    """A counting context on one GPU (stc_create / stc_destroy)."""

    def __init__(self, device=0, model_dir=None):
        L = lib()
        h = C.c_void_p()
        rc = L.stc_create(device, model_dir.encode() if model_dir else None, C.byref(h))
        _check(rc)
        self._h = h
        self.model_status = L.stc_model_status(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().stc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self):
        m = lib().stc_last_error(self._h)
        return m.decode() if m else ""

    def _opts(self, dispatch, repeat, classes):
        d = DISPATCH[dispatch] if isinstance(dispatch, str) else int(dispatch)
        return Opts(d, int(repeat), 1 if classes else 0)

    def count(self, g=None, u=None, v=None, sign=None, n=None, dispatch="learned",
              repeat=1, classes=False):
        """Return (balanced, unbalanced) for a host edge list."""
        gr = _as_graph(g, u, v, sign, n)
        e = gr.edges()
        out = Counts()
        _check(lib().stc_count(self._h, C.byref(e), C.byref(self._opts(dispatch, repeat, classes)),
                               C.byref(out)), self._h)
        return int(out.balanced), int(out.unbalanced)

    def count_device(self, n, m, u_ptr, v_ptr, sign_ptr, dispatch="learned", repeat=1,
                     classes=False):
        """Count a graph whose u / v / sign arrays already live in device memory."""
        e = Edges(n, m, u_ptr, v_ptr, sign_ptr)
        out = Counts()
        _check(lib().stc_count_device(self._h, C.byref(e),
                                      C.byref(self._opts(dispatch, repeat, classes)),
                                      C.byref(out)), self._h)
        return int(out.balanced), int(out.unbalanced)

    def prepare(self, g, device_input=False):
        e = g.edges() if isinstance(g, Graph) else g
        _check(lib().stc_prepare(self._h, C.byref(e), 1 if device_input else 0), self._h)

    def count_prepared(self, dispatch="learned", repeat=1, classes=False):
        out = Counts()
        _check(lib().stc_count_prepared(self._h, C.byref(self._opts(dispatch, repeat, classes)),
                                        C.byref(out)), self._h)
        return int(out.balanced), int(out.unbalanced)

    def count_prepared_static(self, dispatch="learned", repeat=1, classes=False):
        """stc_count_prepared_static: the same count with the counting buffers
        in the library's static device array (no device allocation)."""
        out = Counts()
        _check(lib().stc_count_prepared_static(
            self._h, C.byref(self._opts(dispatch, repeat, classes)), C.byref(out)), self._h)
        return int(out.balanced), int(out.unbalanced)

    def count_per_vertex(self, dispatch="learned", repeat=1):
        """Per-vertex four-class counts of the prepared graph: an (n, 4) uint64
        array, row v = T0..T3 triangles whose lowest-ranked vertex is v."""
        n = self.stats()["n"]
        cnt = np.zeros((max(n, 1), 4), np.uint64)
        out = Counts()
        _check(lib().stc_count_per_vertex(self._h, C.byref(self._opts(dispatch, repeat, True)),
                                          cnt.ctypes.data, C.byref(out)), self._h)
        return cnt[:n]

    def classes(self):
        c = Classes()
        _check(lib().stc_last_classes(self._h, C.byref(c)), self._h)
        return {"t0": int(c.t0), "t1": int(c.t1), "t2": int(c.t2), "t3": int(c.t3)}

    def stats(self):
        s = Stats()
        _check(lib().stc_get_stats(self._h, C.byref(s)), self._h)
        return s.as_dict()

    # ---- dispatcher training / tuning hooks (include/stc_train.h) ----
    def bucket_histogram(self):
        out = np.zeros(256, np.uint32)
        _check(lib().stc_bucket_histogram(self._h, out.ctypes.data), self._h)
        return out

    def count_with_table(self, table, repeat=1):
        t = np.ascontiguousarray(table, np.uint8)
        assert t.shape == (256,)
        out = Counts()
        _check(lib().stc_count_with_table(self._h, t.ctypes.data, repeat, C.byref(out)), self._h)
        return int(out.balanced), int(out.unbalanced)

    def train_dispatcher(self, feat, cost, out_path, iters=3000, lr=0.01, seed=1):
        f = np.ascontiguousarray(feat, np.float32)
        c = np.ascontiguousarray(cost, np.float32)
        assert f.shape[1] == 7 and c.shape[1] == 4 and f.shape[0] == c.shape[0]
        _check(lib().stc_train_dispatcher(self._h, f.ctypes.data, c.ctypes.data, f.shape[0],
                                          iters, lr, seed, out_path.encode()), self._h)

    def last_table(self):
        """The strategy per bucket (TABLE codes) the last count's dispatcher chose."""
        out = np.zeros(256, np.uint8)
        _check(lib().stc_last_table(self._h, out.ctypes.data), self._h)
        return out

    def csr_arrays(self):
        """The prepared graph's device CSR in rank space (stc_get_csr): off,
        col ((rank << 1) | neg), perm (rank -> vertex id), tin_off, tin ((m, 2):
        suffix start, (suffix length << 1) | neg), dcls, smask."""
        st = self.stats()
        n, m = st["n"], st["m_oriented"]
        a = {"off": np.zeros(n + 1, np.uint32), "col": np.zeros(max(m, 1), np.uint32),
             "perm": np.zeros(max(n, 1), np.uint32), "tin_off": np.zeros(n + 1, np.uint32),
             "tin": np.zeros((max(m, 1), 2), np.uint32), "dcls": np.zeros(max(n, 1), np.uint8),
             "smask": np.zeros(max(n, 1), np.uint32)}
        _check(lib().stc_get_csr(self._h, *[a[k].ctypes.data for k in
                                            ("off", "col", "perm", "tin_off", "tin", "dcls",
                                             "smask")]), self._h)
        a["col"], a["tin"] = a["col"][:m], a["tin"][:m]
        a["perm"], a["dcls"], a["smask"] = a["perm"][:n], a["dcls"][:n], a["smask"][:n]
        return a

    def load_models(self, model_dir=None):
        rc = lib().stc_load_models(self._h, model_dir.encode() if model_dir else None)
        self.model_status = rc
        return rc

    def set_tune(self, text):
        _check(lib().stc_set_tune(self._h, text.encode()), self._h)

    @staticmethod
    def _batch_edges(batch):
        if isinstance(batch, Batch):
            k = batch.k
            return k, (Edges * k)(*[batch.graph_edges(g) for g in range(k)])
        k = len(batch)
        return k, (Edges * k)(*[g.edges() for g in batch])

    def count_batch(self, batch, dispatch="learned", repeat=1):
        """Per-graph (balanced, unbalanced) for a Batch or a list of Graphs."""
        k, arr = self._batch_edges(batch)
        out = (Counts * max(k, 1))()
        _check(lib().stc_count_batch(self._h, arr, k,
                                     C.byref(self._opts(dispatch, repeat, False)), out),
               self._h)
        return [(int(out[i].balanced), int(out[i].unbalanced)) for i in range(k)]

    def prepare_batch(self, batch):
        """Upload a batch and build its union CSR (stc_prepare_batch)."""
        k, arr = self._batch_edges(batch)
        _check(lib().stc_prepare_batch(self._h, arr, k), self._h)
        self._batch_k = k

    def count_batch_prepared(self, dispatch="learned", repeat=1):
        k = self._batch_k
        out = (Counts * max(k, 1))()
        _check(lib().stc_count_batch_prepared(self._h, C.byref(self._opts(dispatch, repeat, False)),
                                              out), self._h)
        return [(int(out[i].balanced), int(out[i].unbalanced)) for i in range(k)]


def bucket_features(bucket, count, n, m, k=1):
    """Dispatcher features of a bucket when k graphs of n vertices and m
    oriented edges in total are counted together (k = 1: one graph)."""
    out = np.zeros(7, np.float32)
    lib().stc_bucket_features(bucket, count, n, m, k, out.ctypes.data)
    return out


class signed_tc:
    """Shape of the reference GPU operator (signed_tc.hxx:17-24,527-542)."""

    class param_t:
        pass

    class result_t:
        def __init__(self):
            self.total_balanced = 0
            self.total_unbalanced = 0

    @staticmethod
    def run(G, param=None, result=None, context=None):
        """Count G; fill result.total_balanced / total_unbalanced; return ms."""
        ctx = context if context is not None else Counter()
        b, u = ctx.count(G)
        if result is not None:
            result.total_balanced = b
            result.total_unbalanced = u
        return ctx.stats()["ms_count"]
