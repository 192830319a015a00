"""ctypes binding of the CPU oracle (liboracle.so, built from stc_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")


class _Edges(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("u", C.POINTER(C.c_uint32)),
                ("v", C.POINTER(C.c_uint32)), ("sign", C.POINTER(C.c_int8))]


class _Csr(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64), ("m_in", C.c_uint64),
                ("offsets", C.POINTER(C.c_uint64)), ("edges", C.POINTER(C.c_uint32)),
                ("signs", C.POINTER(C.c_int8)), ("fdeg", C.POINTER(C.c_uint32)),
                ("rank_", C.POINTER(C.c_uint32)), ("graph_hash", C.c_uint64),
                ("max_fdeg", C.c_uint32)]


_L = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        L.orc_parse.argtypes = [C.c_char_p, C.c_uint64, P(P(_Edges))]
        L.orc_load.argtypes = [C.c_char_p, P(P(_Edges))]
        L.orc_free_edges.argtypes = [P(_Edges)]
        L.orc_free_edges.restype = None
        L.orc_build_csr.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int, P(P(_Csr))]
        L.orc_free_csr.argtypes = [P(_Csr)]
        L.orc_free_csr.restype = None
        for f in ("orc_count_merge", "orc_count_hash"):
            getattr(L, f).argtypes = [P(_Csr), C.c_int, P(C.c_uint64), P(C.c_uint64)]
            getattr(L, f).restype = None
        L.orc_count_classes.argtypes = [P(_Csr), C.c_int, P(C.c_uint64)]
        L.orc_count_classes.restype = None
        L.orc_count_classes_per_vertex.argtypes = [P(_Csr), C.c_int, C.c_void_p]
        L.orc_count_classes_per_vertex.restype = None
        L.orc_count_merge_sample.argtypes = [P(_Csr), C.c_int, C.c_uint32, C.c_uint32,
                                             P(C.c_uint64), P(C.c_double), P(C.c_uint64),
                                             P(C.c_uint64)]
        L.orc_count_merge_sample.restype = C.c_double
        L.orc_time_merge.argtypes = [P(_Csr), C.c_int, P(C.c_uint64), P(C.c_uint64)]
        L.orc_time_merge.restype = C.c_double
        L.orc_time_hash.argtypes = [P(_Csr), C.c_int, P(C.c_double), P(C.c_uint64),
                                    P(C.c_uint64)]
        L.orc_time_hash.restype = C.c_double
        _L = L
    return _L


def parse(text):
    """Reference text-format parse (graph_io.h:72-212): (n, u, v, sign)."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    p = C.POINTER(_Edges)()
    lib().orc_parse(data, len(data), C.byref(p))
    e = p.contents
    m = int(e.m)
    u = np.ctypeslib.as_array(e.u, (m,)).copy() if m else np.zeros(0, np.uint32)
    v = np.ctypeslib.as_array(e.v, (m,)).copy() if m else np.zeros(0, np.uint32)
    s = np.ctypeslib.as_array(e.sign, (m,)).copy() if m else np.zeros(0, np.int8)
    n = int(e.n)
    lib().orc_free_edges(p)
    return n, u, v, s


class Csr:
    """The oracle's forward CSR (graph_io.h:22-30) for one edge list."""

    def __init__(self, n, u, v, sign, threads=0):
        u = np.ascontiguousarray(u, np.uint32)
        v = np.ascontiguousarray(v, np.uint32)
        sign = np.ascontiguousarray(sign, np.int8)
        self._keep = (u, v, sign)
        self._p = C.POINTER(_Csr)()
        lib().orc_build_csr(n, len(u), u.ctypes.data, v.ctypes.data, sign.ctypes.data,
                            threads, C.byref(self._p))
        c = self._p.contents
        self.n, self.m = int(c.n), int(c.m)
        self.graph_hash = int(c.graph_hash)
        self.max_fdeg = int(c.max_fdeg)

    def arrays(self):
        c = self._p.contents
        off = np.ctypeslib.as_array(c.offsets, (self.n + 1,)).copy()
        E = np.ctypeslib.as_array(c.edges, (max(self.m, 1),))[: self.m].copy()
        S = np.ctypeslib.as_array(c.signs, (max(self.m, 1),))[: self.m].copy()
        R = np.ctypeslib.as_array(c.rank_, (max(self.n, 1),))[: self.n].copy()
        return off, E, S, R

    def merge(self, threads=0):
        b, u = C.c_uint64(), C.c_uint64()
        lib().orc_count_merge(self._p, threads, C.byref(b), C.byref(u))
        return int(b.value), int(u.value)

    def hash(self, threads=0):
        b, u = C.c_uint64(), C.c_uint64()
        lib().orc_count_hash(self._p, threads, C.byref(b), C.byref(u))
        return int(b.value), int(u.value)

    def classes(self, threads=0):
        t = (C.c_uint64 * 4)()
        lib().orc_count_classes(self._p, threads, t)
        return {"t0": int(t[0]), "t1": int(t[1]), "t2": int(t[2]), "t3": int(t[3])}

    def classes_per_vertex(self, threads=0):
        out = np.zeros(4 * (self.n + 1), np.uint64)
        lib().orc_count_classes_per_vertex(self._p, threads, out.ctypes.data)
        return out[: 4 * self.n].reshape(self.n, 4)

    def time_merge(self, threads=0):
        b, u = C.c_uint64(), C.c_uint64()
        t = lib().orc_time_merge(self._p, threads, C.byref(b), C.byref(u))
        return t, int(b.value), int(u.value)

    def time_hash(self, threads=0):
        """(count seconds, table build seconds, balanced, unbalanced) of the
        hash path."""
        bs, b, u = C.c_double(), C.c_uint64(), C.c_uint64()
        t = lib().orc_time_hash(self._p, threads, C.byref(bs), C.byref(b), C.byref(u))
        return t, float(bs.value), int(b.value), int(u.value)

    def merge_sample(self, stride, phase=0, threads=0):
        ne, by, b, u = C.c_uint64(), C.c_double(), C.c_uint64(), C.c_uint64()
        t = lib().orc_count_merge_sample(self._p, threads, stride, phase, C.byref(ne),
                                         C.byref(by), C.byref(b), C.byref(u))
        return {"seconds": t, "edges": int(ne.value), "bytes": float(by.value),
                "balanced": int(b.value), "unbalanced": int(u.value)}

    def __del__(self):
        if getattr(self, "_p", None) is not None and _L is not None:
            _L.orc_free_csr(self._p)
            self._p = None


def reference_binary(name):
    """Path of a reference program built by `make -C oracle ref`, or None."""
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None
