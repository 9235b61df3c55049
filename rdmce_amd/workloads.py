"""Named synthetic workloads: a graph in degeneracy order plus the edge roots
that are enumerated.

A "core-bounded" workload enumerates every maximal clique whose two lowest
vertices in degeneracy order (the edge root v0 < v1) have core(v0) <= c.
Canonical edge roots are sorted by v0, and v0 ascends with the core number
in a degeneracy order, so these roots are a prefix [0, end) of the canonical
root list.  The graphs below have dense cores whose clique counts grow by
billions per core level near the top (full enumeration is out of reach for
any implementation), so the bound keeps a run to seconds while keeping the
heavy-tailed per-root costs that the schedules are about.

  rmat20-core<c>  R-MAT scale 20, edge factor 16, (0.57,0.19,0.19,0.05),
                  seed 2, symmetrised, deduplicated, no self loops.
  rmat24-core<c>  R-MAT scale 24, edge factor 16, same skew, seed 5.
  cl4m-core<c>    Chung-Lu power law, exponent 2.1, n = 4,000,000, mean
                  degree 20, plus 2,000 planted communities of size uniform
                  in [50, 200] with intra-community edge probability 0.6,
                  seed 4.
  er2048          Erdos-Renyi G(2048, 0.008), seed 1, all roots.

A batch workload is a list of independent graphs enumerated together:

  batch10k-e<x>   The graphs i in [0, 10000) of the family G(n, p), n uniform
                  in [32, 1024], p uniform in [0.05, 0.5], seed 3, whose
                  closed-form expected number of maximal cliques is <= x
                  (e.g. batch10k-e1e6: 6,315 graphs, ~8.5e8 cliques; the
                  whole family expects ~9.5e11).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib
from .graph import Graph, OrderType, batch_graph_params


@dataclasses.dataclass
class Workload:
    name: str
    graph: Graph
    n_roots: int           # canonical edge roots [0, n_roots) are enumerated
    desc: dict

    def plain_order(self) -> np.ndarray:
        """Plain dispatch order: canonical roots in decreasing id (high core
        first), as the reference hands out edge slots
        (BkPivotBitBalance.cuh:429-431)."""
        return np.arange(self.n_roots, dtype=np.uint64)[::-1].copy()


def _core_prefix(g: Graph, core_max: int) -> int:
    core = g.GetCoreNumbers()
    roots = g.roots(_lib.ROOT_EDGE)
    v0 = (roots >> np.uint64(32)).astype(np.int64)
    return int(np.searchsorted(core[v0], core_max, side="right"))


def _split(name: str, prefix: str) -> int:
    return int(name[len(prefix):])


GRAPHS = {
    "rmat20": ("R-MAT scale 20, ef 16, (0.57,0.19,0.19,0.05), seed 2, degeneracy order",
               lambda: Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=2)),
    "rmat24": ("R-MAT scale 24, ef 16, (0.57,0.19,0.19,0.05), seed 5, degeneracy order",
               lambda: Graph.rmat(24, 16, 0.57, 0.19, 0.19, seed=5)),
    "cl4m": ("Chung-Lu gamma 2.1, n 4e6, mean degree 20, 2000 planted communities "
             "(size U[50,200], p_in 0.6), seed 4, degeneracy order",
             lambda: Graph.chung_lu_planted(4_000_000, 2.1, 20.0, 2000, 50, 200, 0.6, 4)),
}


def load(name: str) -> Workload:
    """Builds workload `name` (host generation + degeneracy order)."""
    if name == "er2048":
        g = Graph.erdos_renyi(2048, 0.008, 1)
        g.SortByOrder(OrderType.DEG)
        m = g.GetNumEdges()
        return Workload(name, g, m, {"workload": name, "graph": "G(2048, 0.008) seed 1", "roots": m})
    for key, (text, make) in GRAPHS.items():
        prefix = key + "-core"
        if name.startswith(prefix):
            core_max = _split(name, prefix)
            g = make()
            g.SortByOrder(OrderType.DEG)
            end = _core_prefix(g, core_max)
            desc = {"workload": name, "graph": text, "n": g.GetNumVertices(), "m": g.GetNumEdges(),
                    "degeneracy": g.GetDegeneracy(), "roots": end,
                    "root_rule": f"edge roots (v0<v1) with core(v0) <= {core_max}"}
            return Workload(name, g, end, desc)
    raise ValueError(f"unknown workload {name!r}")


def expected_er_cliques(n: int, p: float) -> float:
    """Expected number of maximal cliques of size >= 2 in G(n, p):
    sum_k C(n,k) p^C(k,2) (1 - p^k)^(n-k), summed in log space."""
    if p <= 0.0:
        return 0.0
    if p >= 1.0:
        return 1.0
    tot = 0.0
    lp = math.log(p)
    for k in range(2, n + 1):
        pk = p ** k
        if pk >= 1.0:
            continue
        lt = (math.lgamma(n + 1) - math.lgamma(k + 1) - math.lgamma(n - k + 1) + math.comb(k, 2) * lp
              + (n - k) * math.log1p(-pk))
        if lt > -700:
            tot += math.exp(lt)
        elif k > n * p + 2:
            break
    return tot


@dataclasses.dataclass
class BatchWorkload:
    name: str
    graphs: list           # Graph objects in degeneracy order
    index: list            # their index in the generated family
    desc: dict


def load_batch(name: str) -> BatchWorkload:
    prefix = "batch10k-e"
    if not name.startswith(prefix):
        raise ValueError(f"unknown batch workload {name!r}")
    cap = float(name[len(prefix):])
    graphs, index = [], []
    for i in range(10000):
        n, p, s = batch_graph_params(3, i, 32, 1024, 0.05, 0.5)
        if expected_er_cliques(n, p) > cap:
            continue
        g = Graph.erdos_renyi(n, p, s)
        g.SortByOrder(OrderType.DEG)
        graphs.append(g)
        index.append(i)
    desc = {"workload": name, "graph": "G(n, p), n ~ U[32,1024], p ~ U[0.05,0.5], seed 3, degeneracy order",
            "family": 10000, "graphs": len(graphs), "rule": f"expected maximal cliques <= {cap:g}"}
    return BatchWorkload(name, graphs, index, desc)


def batch_digest(counts, hashes) -> int:
    """Order-dependent digest of per-graph (count, hash) pairs (mod 2^64)."""
    m = (1 << 64) - 1
    d = 0
    for i, (c, h) in enumerate(zip(counts, hashes)):
        x = (int(c) * 0x9E3779B97F4A7C15 + int(h) + i) & m
        x ^= x >> 31
        x = (x * 0xBF58476D1CE4E5B9) & m
        x ^= x >> 29
        d = (d * 0x100000001B3 + x) & m
    return d
