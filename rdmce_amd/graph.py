"""Host graph, mirroring the reference's ``class Graph``
(RDMCE/include/graph.h:12-52, RDMCE/src/graph.cc).

The graph lives in librdmce.so as a CSR (u64 row offsets, u32 columns); the
accessors return numpy views of it.  Method names follow the reference so
that a caller of ``Graph`` finds the same operations:

    g = Graph()
    g.LoadFromFile("zachary.txt")     # graph.cc:64
    g.SortByOrder(OrderType.DEG)      # graph.cc:241
    g.ConvertToCsr()                  # graph.cc:216 (a no-op: always CSR)
    g.GetNumVertices(), g.GetNumEdges(), g.GetMaxDegree(), g.GetDegeneracy()
"""
from __future__ import annotations

import ctypes
import enum
from ctypes import POINTER, byref, c_uint32, c_uint64, c_void_p

import numpy as np

from . import _lib
from ._lib import check, lib


class OrderType(enum.IntEnum):
    """RDMCE/include/graph.h:4-10."""
    UNCHANGED = _lib.ORDER_UNCHANGED
    ASC = _lib.ORDER_ASC
    DES = _lib.ORDER_DES
    DEG = _lib.ORDER_DEG


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(POINTER(ctype))


class Graph:
    """Owns one ``mce_graph`` handle."""

    def __init__(self, handle: int | None = None, name: str = ""):
        self._h = c_void_p(handle) if handle else None
        self.name = name

    # -- lifetime ---------------------------------------------------------
    def __del__(self):
        self.free()

    def free(self) -> None:
        if getattr(self, "_h", None):
            lib().mce_graph_free(self._h)
            self._h = None

    @property
    def handle(self) -> c_void_p:
        if not self._h:
            raise ValueError("graph is empty: load or generate it first")
        return self._h

    def _replace(self, h: c_void_p) -> None:
        self.free()
        self._h = h

    # -- loaders (graph.cc:64-213) ------------------------------------------
    def LoadFromFile(self, path: str) -> None:
        h = c_void_p()
        check(lib().mce_graph_load_coo_text(path.encode(), byref(h)))
        self._replace(h)
        self.name = path

    def LoadFromBin(self, path: str) -> None:
        h = c_void_p()
        check(lib().mce_graph_load_bin(path.encode(), byref(h)))
        self._replace(h)
        self.name = path

    def StoreIntoBin(self, path: str) -> str:
        """Writes the .bin next to ``path`` with its extension replaced
        (graph.cc:154-156); returns the file written."""
        out = bin_path_for(path)
        check(lib().mce_graph_store_bin(self.handle, out.encode()))
        return out

    def StoreIntoBinExact(self, path: str) -> None:
        check(lib().mce_graph_store_bin(self.handle, path.encode()))

    def WriteCooText(self, path: str) -> None:
        check(lib().mce_graph_write_coo_text(self.handle, path.encode()))

    @classmethod
    def from_pairs(cls, u, v) -> "Graph":
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        if u.shape != v.shape:
            raise ValueError("endpoint arrays differ in length")
        h = c_void_p()
        check(lib().mce_graph_from_pairs(_ptr(u, c_uint32), _ptr(v, c_uint32), u.size, byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, edges) -> "Graph":
        e = np.asarray(list(edges) if not isinstance(edges, np.ndarray) else edges, dtype=np.int64)
        if e.size == 0:
            e = e.reshape(0, 2)
        return cls.from_pairs(e[:, 0], e[:, 1])

    @classmethod
    def from_csr(cls, rowoff, col, labels=None, degeneracy: int = 0) -> "Graph":
        rowoff = np.ascontiguousarray(rowoff, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint32)
        h = c_void_p()
        check(lib().mce_graph_from_csr(rowoff.size - 1, _ptr(rowoff, c_uint64), _ptr(col, c_uint32),
                                       None if lab is None else _ptr(lab, c_uint32), degeneracy, byref(h)))
        return cls(h.value)

    # -- generators ---------------------------------------------------------
    @classmethod
    def erdos_renyi(cls, n: int, p: float, seed: int) -> "Graph":
        h = c_void_p()
        check(lib().mce_gen_er(n, p, seed, byref(h)))
        return cls(h.value, f"er_{n}_{p}_{seed}")

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 0) -> "Graph":
        h = c_void_p()
        check(lib().mce_gen_rmat(scale, edge_factor, a, b, c, seed, byref(h)))
        return cls(h.value, f"rmat{scale}_ef{edge_factor}_s{seed}")

    @classmethod
    def chung_lu_planted(cls, n: int, gamma: float = 2.1, mean_degree: float = 20.0, n_comm: int = 2000,
                         cmin: int = 50, cmax: int = 200, p_in: float = 0.6, seed: int = 4) -> "Graph":
        h = c_void_p()
        check(lib().mce_gen_chung_lu_planted(n, gamma, mean_degree, n_comm, cmin, cmax, p_in, seed, byref(h)))
        return cls(h.value, f"chunglu_{n}_s{seed}")

    # -- ordering (graph.cc:241-306) -----------------------------------------
    def SortByOrder(self, order: OrderType | int = OrderType.DEG) -> None:
        check(lib().mce_graph_order(self.handle, int(order)))

    def ConvertToCsr(self) -> None:
        """graph.cc:216-236.  The graph is always held as CSR."""

    # -- accessors (graph.h:38-46) --------------------------------------------
    def info(self) -> _lib.MceGraphInfo:
        i = _lib.MceGraphInfo()
        check(lib().mce_graph_get_info(self.handle, byref(i)))
        return i

    def GetNumVertices(self) -> int:
        return int(self.info().n)

    def GetNumEdges(self) -> int:
        return int(self.info().m)

    def GetMaxDegree(self) -> int:
        return int(self.info().max_degree)

    def GetDegeneracy(self) -> int:
        return int(self.info().degeneracy)

    def _csr_ptrs(self):
        ro = POINTER(c_uint64)()
        co = POINTER(c_uint32)()
        lb = POINTER(c_uint32)()
        check(lib().mce_graph_get_csr(self.handle, byref(ro), byref(co), byref(lb)))
        return ro, co, lb

    def GetRowOffset(self) -> np.ndarray:
        """Row offsets (n + 1, uint64; the reference uses u32, graph.h:22)."""
        n = self.GetNumVertices()
        ro, _, _ = self._csr_ptrs()
        return np.ctypeslib.as_array(ro, shape=(n + 1,)).copy()

    def GetColIdx(self) -> np.ndarray:
        nnz = 2 * self.GetNumEdges()
        _, co, _ = self._csr_ptrs()
        if nnz == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(co, shape=(nnz,)).copy()

    def GetLabels(self) -> np.ndarray:
        n = self.GetNumVertices()
        _, _, lb = self._csr_ptrs()
        if n == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(lb, shape=(n,)).copy()

    def GetCoreNumbers(self) -> np.ndarray:
        n = self.GetNumVertices()
        p = POINTER(c_uint32)()
        check(lib().mce_graph_get_core(self.handle, byref(p)))
        if n == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def N(self, v: int) -> np.ndarray:
        ro = self.GetRowOffset()
        return self.GetColIdx()[ro[v]:ro[v + 1]]

    def roots(self, root_mode: int = _lib.ROOT_EDGE) -> np.ndarray:
        """Canonical root descriptors (see mce_plan_roots)."""
        cnt = c_uint64()
        check(lib().mce_plan_roots(self.handle, root_mode, None, 0, byref(cnt)))
        out = np.zeros(max(cnt.value, 1), dtype=np.uint64)
        check(lib().mce_plan_roots(self.handle, root_mode, _ptr(out, c_uint64), out.size, byref(cnt)))
        return out[:cnt.value]


def bin_path_for(path: str) -> str:
    buf = ctypes.create_string_buffer(len(path.encode()) + 16)
    check(lib().mce_graph_bin_path(path.encode(), buf, len(buf)))
    return buf.value.decode()


def batch_graph_params(seed: int, index: int, nmin: int = 32, nmax: int = 1024, pmin: float = 0.05,
                       pmax: float = 0.5):
    n = c_uint32()
    p = ctypes.c_double()
    s = c_uint64()
    check(lib().mce_gen_batch_params(seed, index, nmin, nmax, pmin, pmax, byref(n), byref(p), byref(s)))
    return int(n.value), float(p.value), int(s.value)
