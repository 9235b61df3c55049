"""Learned root scheduling.

An FP32 MLP (8 -> 32 -> 32 -> 1, ReLU; HIP kernels in librdmce.so) predicts
log(1 + cycles) of every edge root from structural features computed on the
GPU.  Labels come from a plain run with per-root statistics (the cycles the
waves spent on each root, summed over waves when it was shared).  The
schedule dispatches roots in decreasing predicted cost (longest first) and
flags the predicted heaviest for an eager split: their top level is exposed
to other waves as soon as it is known.  Ordering and splitting never change
results; with a model that predicts garbage (non-finite outputs) the plain
order is used.
"""
from __future__ import annotations

import ctypes
import dataclasses
from ctypes import POINTER, byref, c_float, c_uint64, c_void_p

import numpy as np

from ._lib import check, lib

NFEAT = 8


def _fptr(a):
    return a.ctypes.data_as(POINTER(c_float))


def _setup():
    L = lib()
    if getattr(L, "_model_sigs", False):
        return L
    G = c_void_p
    sigs = {
        "mce_root_features": ([G, ctypes.c_int32, POINTER(c_float), c_uint64, POINTER(c_uint64)], ctypes.c_int32),
        "mce_cost_model_num_params": ([], ctypes.c_uint32),
        "mce_cost_model_num_trainable": ([], ctypes.c_uint32),
        "mce_cost_model_create": ([c_uint64, POINTER(c_void_p)], ctypes.c_int32),
        "mce_cost_model_get_params": ([c_void_p, POINTER(c_float), c_uint64], ctypes.c_int32),
        "mce_cost_model_set_params": ([c_void_p, POINTER(c_float), c_uint64], ctypes.c_int32),
        "mce_cost_model_fit_normalizer": ([c_void_p, POINTER(c_float), c_uint64], ctypes.c_int32),
        "mce_cost_model_gradient": ([c_void_p, ctypes.c_int32, POINTER(c_float), POINTER(c_float), c_uint64,
                                     POINTER(c_float), POINTER(c_float)], ctypes.c_int32),
        "mce_cost_model_train": ([c_void_p, ctypes.c_int32, POINTER(c_float), POINTER(c_float), c_uint64,
                                  ctypes.c_uint32, ctypes.c_uint32, c_float, c_uint64, POINTER(c_float)],
                                 ctypes.c_int32),
        "mce_cost_model_predict": ([c_void_p, ctypes.c_int32, POINTER(c_float), c_uint64, POINTER(c_float)],
                                   ctypes.c_int32),
        "mce_cost_model_free": ([c_void_p], None),
    }
    for name, (args, ret) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ret
    L._model_sigs = True
    return L


def root_features(graph, device: int = 0) -> np.ndarray:
    """[m, 8] float32 features of the canonical edge roots (GPU)."""
    L = _setup()
    n = c_uint64()
    check(L.mce_root_features(graph.handle, device, None, 0, byref(n)))
    out = np.zeros((max(n.value, 1), NFEAT), dtype=np.float32)
    check(L.mce_root_features(graph.handle, device, _fptr(out), out.size, byref(n)))
    return out[:n.value]


class CostModel:
    def __init__(self, seed: int = 0):
        self.L = _setup()
        self._h = c_void_p()
        check(self.L.mce_cost_model_create(seed, byref(self._h)))

    @property
    def n_params(self) -> int:
        return int(self.L.mce_cost_model_num_params())

    @property
    def n_trainable(self) -> int:
        return int(self.L.mce_cost_model_num_trainable())

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        check(self.L.mce_cost_model_get_params(self._h, _fptr(p), p.size))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(self.L.mce_cost_model_set_params(self._h, _fptr(p), p.size))

    def fit_normalizer(self, X) -> None:
        X = np.ascontiguousarray(X, dtype=np.float32)
        check(self.L.mce_cost_model_fit_normalizer(self._h, _fptr(X), X.shape[0]))

    def gradient(self, X, Y, device: int = 0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.n_trainable, dtype=np.float32)
        loss = c_float()
        check(self.L.mce_cost_model_gradient(self._h, device, _fptr(X), _fptr(Y), X.shape[0], _fptr(g), byref(loss)))
        return g, float(loss.value)

    def train(self, X, Y, epochs: int = 20, batch: int = 1024, lr: float = 3e-3, seed: int = 0,
              device: int = 0) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        loss = c_float()
        check(self.L.mce_cost_model_train(self._h, device, _fptr(X), _fptr(Y), X.shape[0], epochs, batch, lr, seed,
                                          byref(loss)))
        return float(loss.value)

    def predict(self, X, device: int = 0) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(max(X.shape[0], 1), dtype=np.float32)
        check(self.L.mce_cost_model_predict(self._h, device, _fptr(X), X.shape[0], _fptr(out)))
        return out[:X.shape[0]]

    def free(self):
        if getattr(self, "_h", None):
            self.L.mce_cost_model_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


@dataclasses.dataclass
class Schedule:
    order: np.ndarray      # canonical root ids in dispatch order
    eager: np.ndarray      # uint8 per dispatch position
    learned: bool          # False: fell back to the plain order
    runs: np.ndarray | None = None  # packed search: run boundaries in `order` (one first vertex per run)


def plain_schedule(root_ids) -> Schedule:
    """Plain method: roots in decreasing canonical id (high core first),
    BkPivotBitBalance.cuh:429-431."""
    ids = np.sort(np.asarray(root_ids, dtype=np.uint64))[::-1].copy()
    return Schedule(ids, np.zeros(ids.size, dtype=np.uint8), False)


def learned_schedule(pred, root_ids, eager_top: int = 0, heavy_fraction: float = 1.0) -> Schedule:
    """Order of `root_ids` (canonical ids) from the per-root predictions
    `pred` (indexed by canonical id): the predicted-heaviest `heavy_fraction`
    of the roots first, longest predicted first, then the others in the
    plain order (1.0: every root by prediction).  Keeping the bulk in the
    plain order keeps its locality in the CSR (neighbouring roots share
    adjacency rows) while the heavy tail starts early.  The `eager_top`
    heaviest are split at launch.  Falls back to the plain order when a
    prediction is not finite."""
    root_ids = np.asarray(root_ids, dtype=np.uint64)
    p = np.asarray(pred, dtype=np.float64)[root_ids.astype(np.int64)]
    if p.size and not np.all(np.isfinite(p)):
        return plain_schedule(root_ids)
    # stable: equal predictions keep the plain (decreasing id) order
    base = np.sort(root_ids)[::-1]
    pb = np.asarray(pred, dtype=np.float64)[base.astype(np.int64)]
    by_pred = np.argsort(-pb, kind="stable")
    k = int(round(min(max(heavy_fraction, 0.0), 1.0) * base.size))
    if k >= base.size:
        order = base[by_pred].copy()
    else:
        heavy = by_pred[:k]
        rest = np.ones(base.size, dtype=bool)
        rest[heavy] = False
        order = np.concatenate([base[heavy], base[rest]])
    eager = np.zeros(order.size, dtype=np.uint8)
    eager[:min(eager_top, order.size)] = 1
    return Schedule(order, eager, True)


def learned_split(pred, root_ids, heavy_fraction: float) -> Schedule:
    """Schedule for the packed search: the roots stay in the plain order
    (runs of edge roots of one first vertex share that vertex's bitmap), and
    the predicted-heaviest `heavy_fraction` of them are flagged, which sends
    them to the one-root-per-wave kernel that splits them across waves - a
    packed wave searches a root alone, so one heavy root would hold it to the
    end of the launch.  Falls back to the plain order (no flags) when a
    prediction is not finite."""
    base = np.sort(np.asarray(root_ids, dtype=np.uint64))[::-1].copy()
    p = np.asarray(pred, dtype=np.float64)[base.astype(np.int64)]
    eager = np.zeros(base.size, dtype=np.uint8)
    if p.size and not np.all(np.isfinite(p)):
        return Schedule(base, eager, False)
    k = int(round(min(max(heavy_fraction, 0.0), 1.0) * base.size))
    if k:
        eager[np.argsort(-p, kind="stable")[:k]] = 1
    return Schedule(base, eager, True)


def learned_group_schedule(pred, root_ids, first_vertex, waves: int, chunk_share: float = 0.25,
                           route_share: float = 1.0) -> Schedule:
    """Schedule for the packed search from per-root predicted costs.

    The roots are taken in runs that share their first vertex v0 (the packed
    search builds v0's bitmap once per run; `first_vertex` gives v0 for
    every canonical root id).  With c = the predicted cost of a root and
    S = sum(c) / waves (one wave's share of the whole run):
      * a run predicted to cost more than chunk_share * S is split into
        chunks of about that cost, which different waves take;
      * runs and chunks are dispatched in decreasing predicted cost, so the
        ones that could end last start first;
      * a root predicted to cost more than route_share * S is flagged: the
        packed search hands it to the one-root-per-wave kernel, which splits
        it across waves (a packed wave searches a root alone).
    Falls back to the plain order when a prediction is not finite."""
    base = np.sort(np.asarray(root_ids, dtype=np.uint64))[::-1].copy()
    logc = np.asarray(pred, dtype=np.float64)[base.astype(np.int64)]
    if logc.size == 0:
        return Schedule(base, np.zeros(0, dtype=np.uint8), False)
    if not np.all(np.isfinite(logc)):
        return plain_schedule(root_ids)
    c = np.expm1(np.clip(logc, 0.0, 60.0))
    share = c.sum() / max(int(waves), 1)
    v0 = np.asarray(first_vertex)[base.astype(np.int64)]
    new_run = np.empty(base.size, dtype=bool)
    new_run[0] = True
    new_run[1:] = v0[1:] != v0[:-1]
    run = np.cumsum(new_run) - 1
    cum = np.cumsum(c)
    run_start_cum = (cum - c)[new_run][run]  # cost before the run's first root
    chunk_cost = max(chunk_share * share, 1e-30)
    chunk_in_run = np.floor((cum - c - run_start_cum) / chunk_cost).astype(np.int64)
    piece = np.cumsum(new_run | np.r_[False, chunk_in_run[1:] != chunk_in_run[:-1]]) - 1
    piece_cost = np.bincount(piece, weights=c)
    rank = np.argsort(np.argsort(-piece_cost, kind="stable"), kind="stable")
    perm = np.lexsort((np.arange(base.size), rank[piece]))
    order = base[perm]
    pieces = piece[perm]
    runs = np.r_[0, np.flatnonzero(pieces[1:] != pieces[:-1]) + 1, order.size].astype(np.uint64)
    eager = (c[perm] > route_share * share).astype(np.uint8)
    return Schedule(order, eager, True, runs)


def training_set(features, root_cycles, root_ids, max_samples: int = 1 << 20, seed: int = 0):
    """(X, Y) from a plain run's per-root cycles: Y = log1p(cycles)."""
    rng = np.random.default_rng(seed)
    ids = np.asarray(root_ids, dtype=np.int64)
    if ids.size > max_samples:
        ids = rng.choice(ids, size=max_samples, replace=False)
    X = np.ascontiguousarray(features[ids], dtype=np.float32)
    Y = np.log1p(np.asarray(root_cycles, dtype=np.float64)[ids]).astype(np.float32)
    return X, Y


def spearman(a, b) -> float:
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    d = np.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float((ra * rb).sum() / d) if d else 0.0


@dataclasses.dataclass
class FitReport:
    sample_roots: int        # roots labelled by the sample run
    sample_kernel_ms: float  # the sample run's kernel time
    loss: float              # final training loss (mean squared error of log1p(cycles))
    seconds: float           # wall time of labelling + features + training + prediction


def fit_and_predict(graph, n_roots: int, stride: int = 16, epochs: int = 30, batch: int = 1024,
                    lr: float = 3e-3, seed: int = 1, device: int = 0, **plan_kw):
    """Learns per-root costs of the edge roots [0, n_roots) of `graph`:
    every `stride`-th root is labelled with its measured cost (cycles) by
    one plain run of that sample, a CostModel is trained on those labels by
    the GPU kernels, and its predictions for all n_roots roots are returned
    with the model and a FitReport."""
    import time

    from .solver import Plan

    t0 = time.perf_counter()
    sample = np.arange(0, n_roots, stride, dtype=np.uint64)
    plan = Plan(graph, root_order=sample[::-1].copy(), want_root_stats=True, device=device, **plan_kw)
    try:
        out = plan.run()
    finally:
        plan.free()
    F = root_features(graph, device)[:n_roots]
    X, Y = training_set(F, out.root_cycles, sample.astype(np.int64), seed=seed)
    model = CostModel(seed)
    model.fit_normalizer(X)
    loss = model.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed, device=device)
    pred = model.predict(F, device)
    rep = FitReport(int(sample.size), float(out.kernel_ms), float(loss), time.perf_counter() - t0)
    return pred, model, rep
