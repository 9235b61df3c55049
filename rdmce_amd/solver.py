"""Solver front end, mirroring the reference's operator surface:

* ``BkSolverWrapper(graph, device_id) -> int``  (RDMCE/src/kernels/
  BkPivotBitBalance.cuh:611-678): upload, enumerate, free; returns the count.
* ``MceGpuSolver(graph, device_id).Solve()``    (RDMCE/src/mce_gpu.cu:9-22):
  the same, timed, printing ``MC numbers:`` and ``MCE GPU Time:`` exactly as
  the reference does (scripts/analyze.sh:97 reads field 4 of the time line).

Beyond the reference (which only counts) ``solve()`` returns the size
histogram, the clique hash, the optional clique list and per-root search
statistics.  All work runs in librdmce.so's HIP kernels; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time
from ctypes import POINTER, byref, c_uint64, c_void_p

import numpy as np

from . import _lib
from ._lib import MceOptions, MceResult, check, lib
from .graph import Graph


@dataclasses.dataclass
class MceOutput:
    count: int
    hash: int
    hist: np.ndarray            # hist[k] = maximal cliques of size k
    cliques: list | None        # sorted tuples of original labels (want_list)
    root_nodes: np.ndarray | None
    root_cycles: np.ndarray | None
    total_nodes: int
    deferred_roots: int
    shared_tasks: int
    kernel_ms: float
    total_ms: float
    passes: int
    waves: int
    search_bytes: int
    build_bytes: int
    published_levels: int = 0
    owner_tasks: int = 0

    def hist_dict(self) -> dict:
        return {int(k): int(v) for k, v in enumerate(self.hist) if v}


def make_options(device: int = 0, root_mode: int = _lib.ROOT_EDGE, waves_per_cu: int = 0, share_bound: int = 0,
                 want_list: bool = False, want_root_stats: bool = False, list_capacity: int = 0,
                 slot_bytes: int = 0, root_order=None, lds_tier: int = 1, use_graph: bool = False,
                 lds_bytes: int = 0, eager=None, word_path: bool = True, pack: bool = False, pack_slots: int = 0,
                 pack_waves_per_cu: int = 0, runs=None) -> tuple[MceOptions, object]:
    o = MceOptions()
    lib().mce_options_init(byref(o))
    o.device = device
    o.root_mode = root_mode
    o.waves_per_cu = waves_per_cu
    o.share_bound = share_bound
    o.want_list = int(want_list)
    o.want_root_stats = int(want_root_stats)
    o.list_capacity = list_capacity
    o.slot_bytes = slot_bytes
    o.lds_tier = lds_tier
    o.use_graph = int(use_graph)
    o.lds_bytes = lds_bytes
    o.word_path = 0 if word_path else -1
    o.pack = int(pack)
    o.pack_slots = pack_slots
    o.pack_waves_per_cu = pack_waves_per_cu
    keep = []
    if root_order is not None:
        ro = np.ascontiguousarray(root_order, dtype=np.uint64)
        keep.append(ro)
        o.schedule = _lib.SCHED_CUSTOM
        o.roots = ro.ctypes.data_as(POINTER(c_uint64))
        o.n_roots = ro.size
    if runs is not None:
        ru = np.ascontiguousarray(runs, dtype=np.uint64)
        keep.append(ru)
        o.runs = ru.ctypes.data_as(POINTER(c_uint64))
        o.n_runs = max(ru.size - 1, 0)
    if eager is not None:
        ea = np.ascontiguousarray(eager, dtype=np.uint8)
        keep.append(ea)
        o.eager = ea.ctypes.data_as(POINTER(ctypes.c_uint8))
    return o, keep


def _convert(r: MceResult, want_list: bool, want_root_stats: bool) -> MceOutput:
    hist = np.ctypeslib.as_array(r.hist, shape=(r.hist_len,)).copy() if r.hist_len else np.zeros(0, np.uint64)
    cliques = None
    if want_list:
        flat = np.ctypeslib.as_array(r.list, shape=(max(r.list_len, 1),))[:r.list_len].copy()
        cliques = []
        i = 0
        while i < flat.size:
            k = int(flat[i])
            cliques.append(tuple(sorted(int(x) for x in flat[i + 1:i + 1 + k])))
            i += k + 1
    rn = rc = None
    if want_root_stats and r.n_roots:
        rn = np.ctypeslib.as_array(r.root_nodes, shape=(r.n_roots,)).copy()
        rc = np.ctypeslib.as_array(r.root_cycles, shape=(r.n_roots,)).copy()
    return MceOutput(int(r.count), int(r.hash), hist, cliques, rn, rc, int(r.total_nodes), int(r.deferred_roots),
                     int(r.shared_tasks), float(r.kernel_ms), float(r.total_ms), int(r.passes), int(r.waves),
                     int(r.search_bytes), int(r.build_bytes), int(r.published_levels), int(r.owner_tasks))


def solve(graph: Graph, **kw) -> MceOutput:
    """One full call (upload + enumerate + free), like BkSolverWrapper."""
    o, keep = make_options(**kw)
    res = POINTER(MceResult)()
    check(lib().mce_solve(graph.handle, byref(o), byref(res)))
    try:
        return _convert(res.contents, bool(o.want_list), bool(o.want_root_stats))
    finally:
        lib().mce_result_free(res)
        del keep


class Plan:
    """Graph, roots and scratch resident in HBM; ``run()`` enumerates only."""

    def __init__(self, graph: Graph, **kw):
        self._o, self._keep = make_options(**kw)
        self._h = c_void_p()
        check(lib().mce_plan_create(graph.handle, byref(self._o), byref(self._h)))

    def run(self) -> MceOutput:
        res = POINTER(MceResult)()
        check(lib().mce_plan_run(self._h, byref(res)))
        try:
            return _convert(res.contents, bool(self._o.want_list), bool(self._o.want_root_stats))
        finally:
            lib().mce_result_free(res)

    def autotune(self, stride: int = 1) -> dict:
        """Picks waves/CU, LDS tier and sharing bound on a sample of the roots
        (every `stride`-th); later runs use the winner."""
        t = _lib.MceTuneResult()
        check(lib().mce_plan_autotune(self._h, stride, byref(t)))
        return _tune_dict(t)

    def free(self) -> None:
        if getattr(self, "_h", None):
            lib().mce_plan_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


def choose_root_mode(graph: Graph, n_vertex_roots: int = 0, stride: int = 64, repeats: int = 2, **plan_kw) -> dict:
    """Root granularity for `graph`: vertex roots (the reference's LEVEL_1)
    or edge roots (LEVEL_2), BkPivotBitBalance.cuh:414-427.

    Times the vertex roots v of every `stride`-th vertex among the first
    `n_vertex_roots` (all vertices when 0) against the edge roots (v, u) of
    the same vertices, with the plan settings `plan_kw` (best of `repeats`
    runs after one untimed run), and returns the faster mode with both
    times.  Both root sets enumerate the maximal cliques whose lowest vertex
    is a sampled v; a differing count or hash raises."""
    n = graph.GetNumVertices() if n_vertex_roots <= 0 else min(n_vertex_roots, graph.GetNumVertices())
    vs = np.arange(0, n, max(stride, 1), dtype=np.uint64)
    v0 = graph.roots(_lib.ROOT_EDGE) >> np.uint64(32)
    lo = np.searchsorted(v0, vs, side="left")
    hi = np.searchsorted(v0, vs, side="right")
    edges = np.concatenate([np.arange(a, b, dtype=np.uint64) for a, b in zip(lo, hi)] or
                           [np.zeros(0, dtype=np.uint64)])
    kw = {k: v for k, v in plan_kw.items() if k not in ("root_mode", "root_order", "eager", "runs")}
    out = {}
    for label, mode, order in (("vertex", _lib.ROOT_VERTEX, vs), ("edge", _lib.ROOT_EDGE, edges)):
        if order.size == 0:
            out[label] = (0.0, 0, 0)
            continue
        plan = Plan(graph, root_mode=mode, root_order=order, **kw)
        try:
            plan.run()
            best = None
            for _ in range(max(repeats, 1)):
                o = plan.run()
                best = o.kernel_ms if best is None else min(best, o.kernel_ms)
            out[label] = (best, o.count, o.hash)
        finally:
            plan.free()
    if out["vertex"][1:] != out["edge"][1:]:
        raise RuntimeError(f"vertex and edge roots of the sample disagree: {out['vertex'][1:]} != {out['edge'][1:]}")
    mode = _lib.ROOT_VERTEX if out["vertex"][0] < out["edge"][0] else _lib.ROOT_EDGE
    return {"root_mode": mode, "vertex_ms": out["vertex"][0], "edge_ms": out["edge"][0],
            "sample_vertices": int(vs.size), "sample_edge_roots": int(edges.size), "sample_cliques": out["edge"][1]}


def _tune_dict(t) -> dict:
    return {"waves_per_cu": t.waves_per_cu, "share_bound": t.share_bound, "lds_tier": t.lds_tier,
            "trials": t.trials, "best_ms": t.best_ms, "base_ms": t.base_ms, "pack_slots": t.pack_slots,
            "pack_waves_per_cu": t.pack_waves_per_cu}


@dataclasses.dataclass
class BatchOutput:
    count: np.ndarray           # per graph
    hash: np.ndarray            # per graph
    hist: np.ndarray            # [graph, size]
    total_count: int
    total_nodes: int
    kernel_ms: float
    total_ms: float
    graph_replay: bool
    passes: int

    def hist_dict(self, g: int) -> dict:
        return {int(k): int(v) for k, v in enumerate(self.hist[g]) if v}


class Batch:
    """Many independent graphs enumerated by one persistent launch
    (optionally replayed as a captured hipGraph: use_graph=True)."""

    def __init__(self, graphs, **kw):
        self._o, self._keep = make_options(**kw)
        arr = (c_void_p * len(graphs))(*[g.handle.value for g in graphs])
        self._h = c_void_p()
        check(lib().mce_batch_create(arr, len(graphs), byref(self._o), byref(self._h)))

    def run(self) -> BatchOutput:
        res = POINTER(_lib.MceBatchResult)()
        check(lib().mce_batch_run(self._h, byref(res)))
        try:
            r = res.contents
            n = int(r.n_graphs)
            cnt = np.ctypeslib.as_array(r.count, shape=(n,)).copy()
            hs = np.ctypeslib.as_array(r.hash, shape=(n,)).copy()
            hist = np.ctypeslib.as_array(r.hist, shape=(n * r.hist_stride,)).copy().reshape(n, r.hist_stride)
            return BatchOutput(cnt, hs, hist, int(r.total_count), int(r.total_nodes), float(r.kernel_ms),
                               float(r.total_ms), bool(r.graph_replay), int(r.passes))
        finally:
            lib().mce_batch_result_free(res)

    def autotune(self, stride: int = 1) -> dict:
        t = _lib.MceTuneResult()
        check(lib().mce_batch_autotune(self._h, stride, byref(t)))
        return _tune_dict(t)

    def free(self) -> None:
        if getattr(self, "_h", None):
            lib().mce_batch_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


def BkSolverWrapper(graph: Graph, device_id: int = 0) -> int:
    """RDMCE/src/kernels/BkPivotBitBalance.cuh:611 - returns the clique count."""
    return solve(graph, device=device_id).count


class MceGpuSolver:
    """RDMCE/src/mce_gpu.cu:9-22."""

    def __init__(self, g: Graph, device_id: int = 0):
        self.g_ = g
        self.device_id_ = device_id
        self.mce_count_ = 0
        self.result: MceOutput | None = None

    def Solve(self) -> int:
        t0 = time.perf_counter()
        self.result = solve(self.g_, device=self.device_id_)
        elapsed = time.perf_counter() - t0
        self.mce_count_ = self.result.count
        print(f"\nMC numbers: {self.mce_count_}")
        print(f"MCE GPU Time: {elapsed}")
        return self.mce_count_


def device_count() -> int:
    return int(lib().mce_device_count())
