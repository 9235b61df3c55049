"""ctypes binding of the CPU oracle (oracle/liboracle.so).  Test
infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only, as the checker."""
from __future__ import annotations

import ctypes
import dataclasses
import os
from ctypes import POINTER, Structure, byref, c_char_p, c_double, c_int, c_uint32, c_uint64

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_MCE = os.path.join(ROOT, "oracle", "_ref", "ref_mce")
HIST_MAX = 4096


class OrcGraph(Structure):
    _fields_ = [("n", c_uint64), ("m", c_uint64), ("max_degree", c_uint64), ("degeneracy", c_uint64),
                ("rowoff", POINTER(c_uint64)), ("col", POINTER(c_uint32)), ("labels", POINTER(c_uint32))]


class OrcResult(Structure):
    _fields_ = [("count", c_uint64), ("hash", c_uint64), ("hist", c_uint64 * HIST_MAX), ("total_nodes", c_uint64),
                ("n_roots", c_uint64), ("root_nodes", POINTER(c_uint64)), ("list_len", c_uint64),
                ("list", POINTER(c_uint32)), ("list_cap", c_uint64), ("overflow", c_int), ("seconds", c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        PG = POINTER(POINTER(OrcGraph))
        PR = POINTER(POINTER(OrcResult))
        L.orc_graph_from_pairs.argtypes = [POINTER(c_uint32), POINTER(c_uint32), c_uint64, PG]
        L.orc_graph_load_coo.argtypes = [c_char_p, PG]
        L.orc_graph_from_csr.argtypes = [c_uint64, POINTER(c_uint64), POINTER(c_uint32), POINTER(c_uint32),
                                         c_uint64, PG]
        L.orc_graph_order_degeneracy.argtypes = [POINTER(OrcGraph)]
        L.orc_graph_free.argtypes = [POINTER(OrcGraph)]
        L.orc_mce.argtypes = [POINTER(OrcGraph), c_int, c_int, c_int, c_uint64, PR]
        L.orc_mce_range.argtypes = [POINTER(OrcGraph), c_int, c_int, c_int, c_uint64, c_uint64, c_uint64, PR]
        L.orc_mce_tomita.argtypes = [POINTER(OrcGraph), PR]
        L.orc_mce_roots.argtypes = [POINTER(OrcGraph), c_int, c_int, POINTER(c_uint64), c_uint64, PR]
        L.orc_result_free.argtypes = [POINTER(OrcResult)]
        L.orc_build_terms.argtypes = [POINTER(OrcGraph), c_uint64, c_uint64, c_int, POINTER(c_uint64)]
        L.orc_splitmix64.argtypes = [c_uint64]
        L.orc_splitmix64.restype = c_uint64
        L.orc_num_roots.argtypes = [POINTER(OrcGraph), c_int]
        L.orc_num_roots.restype = c_uint64
        _lib = L
    return _lib


@dataclasses.dataclass
class OracleOut:
    count: int
    hash: int
    hist: dict
    total_nodes: int
    root_nodes: np.ndarray | None
    cliques: list | None
    seconds: float


class OracleGraph:
    def __init__(self, ptr):
        self.p = ptr

    def __del__(self):
        if getattr(self, "p", None):
            lib().orc_graph_free(self.p)
            self.p = None

    @classmethod
    def from_pairs(cls, u, v):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        p = POINTER(OrcGraph)()
        lib().orc_graph_from_pairs(u.ctypes.data_as(POINTER(c_uint32)), v.ctypes.data_as(POINTER(c_uint32)),
                                   u.size, byref(p))
        return cls(p)

    @classmethod
    def from_edges(cls, edges):
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        return cls.from_pairs(e[:, 0], e[:, 1])

    @classmethod
    def load_coo(cls, path):
        p = POINTER(OrcGraph)()
        if lib().orc_graph_load_coo(path.encode(), byref(p)) != 0:
            raise IOError(path)
        return cls(p)

    @classmethod
    def from_csr(cls, rowoff, col, labels=None, degeneracy=0):
        rowoff = np.ascontiguousarray(rowoff, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint32)
        p = POINTER(OrcGraph)()
        lib().orc_graph_from_csr(rowoff.size - 1, rowoff.ctypes.data_as(POINTER(c_uint64)),
                                 col.ctypes.data_as(POINTER(c_uint32)),
                                 None if lab is None else lab.ctypes.data_as(POINTER(c_uint32)), degeneracy,
                                 byref(p))
        return cls(p)

    def order_degeneracy(self):
        lib().orc_graph_order_degeneracy(self.p)
        return self

    @property
    def n(self):
        return int(self.p.contents.n)

    @property
    def m(self):
        return int(self.p.contents.m)

    @property
    def degeneracy(self):
        return int(self.p.contents.degeneracy)

    @property
    def max_degree(self):
        return int(self.p.contents.max_degree)

    def csr(self):
        g = self.p.contents
        ro = np.ctypeslib.as_array(g.rowoff, shape=(g.n + 1,)).copy()
        co = np.ctypeslib.as_array(g.col, shape=(max(int(ro[-1]), 1),))[:int(ro[-1])].copy()
        lb = np.ctypeslib.as_array(g.labels, shape=(max(g.n, 1),))[:g.n].copy()
        return ro, co, lb

    def num_roots(self, root_mode=2):
        return int(lib().orc_num_roots(self.p, root_mode))

    def mce(self, root_mode=2, threads=0, root_nodes=False, list_cap=0, root_begin=None, root_end=None):
        r = POINTER(OrcResult)()
        if root_begin is None:
            lib().orc_mce(self.p, root_mode, threads, int(root_nodes), list_cap, byref(r))
        else:
            lib().orc_mce_range(self.p, root_mode, threads, int(root_nodes), list_cap, root_begin, root_end,
                                byref(r))
        try:
            return _convert(r.contents, root_nodes, list_cap)
        finally:
            lib().orc_result_free(r)

    def build_terms(self, root_begin=0, root_end=None, threads=0):
        """{roots, deg, pdeg, rows_w, build_bytes} of the edge roots
        [root_begin, root_end): the algorithmic build bytes 4 deg + 4 pdeg +
        8 rows_w (d(v0) + d(v1), the degrees of P0's members, rows x words of
        each root's bitmap)."""
        end = self.p.contents.m if root_end is None else root_end
        out = (c_uint64 * 4)()
        if lib().orc_build_terms(self.p, root_begin, end, threads, out) != 0:
            raise MemoryError("orc_build_terms")
        roots, deg, pdeg, rows_w = (int(v) for v in out)
        return {"roots": roots, "deg": deg, "pdeg": pdeg, "rows_w": rows_w,
                "build_bytes": 4 * deg + 4 * pdeg + 8 * rows_w}

    def mce_roots(self, ids, root_mode=2, threads=0):
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        r = POINTER(OrcResult)()
        if lib().orc_mce_roots(self.p, root_mode, threads, ids.ctypes.data_as(POINTER(c_uint64)), ids.size,
                               byref(r)) != 0:
            raise ValueError("root id out of range")
        try:
            return _convert(r.contents, False, 0)
        finally:
            lib().orc_result_free(r)

    def tomita(self):
        r = POINTER(OrcResult)()
        lib().orc_mce_tomita(self.p, byref(r))
        try:
            return _convert(r.contents, False, 0)
        finally:
            lib().orc_result_free(r)


def _convert(c: OrcResult, root_nodes: bool, list_cap: int) -> OracleOut:
    hist = {k: int(c.hist[k]) for k in range(HIST_MAX) if c.hist[k]}
    rn = None
    if root_nodes and c.n_roots:
        rn = np.ctypeslib.as_array(c.root_nodes, shape=(c.n_roots,)).copy()
    cl = None
    if list_cap:
        if c.overflow:
            raise RuntimeError("oracle clique list overflow")
        flat = np.ctypeslib.as_array(c.list, shape=(max(c.list_len, 1),))[:c.list_len]
        cl = []
        i = 0
        while i < flat.size:
            k = int(flat[i])
            cl.append(tuple(int(x) for x in flat[i + 1:i + 1 + k]))
            i += k + 1
    return OracleOut(int(c.count), int(c.hash), hist, int(c.total_nodes), rn, cl, float(c.seconds))


def splitmix64(x: int) -> int:
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def clique_hash(cliques) -> int:
    """Order-independent hash of a clique list (original labels)."""
    M = (1 << 64) - 1
    h = 0
    for c in cliques:
        key = 0
        for v in c:
            key = (key + splitmix64(int(v))) & M
        h = (h + splitmix64(key)) & M
    return h
