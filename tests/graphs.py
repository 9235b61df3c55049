"""Small test graphs with known maximal-clique structure (closed forms)."""
from __future__ import annotations

import itertools
import math

import numpy as np


def complete(n):
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


def cycle(n):
    return [(i, (i + 1) % n) for i in range(n)]


def star(m):
    return [(0, i) for i in range(1, m + 1)]


def complete_multipartite(sizes):
    part = []
    for idx, s in enumerate(sizes):
        part += [idx] * s
    n = len(part)
    return [(i, j) for i in range(n) for j in range(i + 1, n) if part[i] != part[j]]


def moon_moser(k):
    """K_{3,3,...,3} with k parts: 3^k maximal cliques of size k."""
    return complete_multipartite([3] * k)


def cocktail_party(k):
    """2k vertices, all edges but a perfect matching: 2^k cliques of size k."""
    n = 2 * k
    return [(i, j) for i in range(n) for j in range(i + 1, n) if j != i + k]


def disjoint_cliques(sizes):
    edges, base = [], 0
    for s in sizes:
        edges += [(base + i, base + j) for i in range(s) for j in range(i + 1, s)]
        base += s
    return edges


def closed_forms():
    """(name, edges, expected histogram {size: count})."""
    out = [
        ("K2", complete(2), {2: 1}),
        ("K5", complete(5), {5: 1}),
        ("K70", complete(70), {70: 1}),
        ("K130", complete(130), {130: 1}),
        ("C4", cycle(4), {2: 4}),
        ("C9", cycle(9), {2: 9}),
        ("C3", cycle(3), {3: 1}),
        ("star7", star(7), {2: 7}),
        ("moon_moser4", moon_moser(4), {4: 81}),
        ("moon_moser7", moon_moser(7), {7: 3 ** 7}),
        ("cocktail8", cocktail_party(8), {8: 2 ** 8}),
        ("cocktail20", cocktail_party(20), {20: 2 ** 20}),
        ("multipartite_2_3_4", complete_multipartite([2, 3, 4]), {3: 24}),
        ("multipartite_1_5_9_2", complete_multipartite([1, 5, 9, 2]), {4: 90}),
        ("disjoint_3_4_5_2", disjoint_cliques([3, 4, 5, 2]), {3: 1, 4: 1, 5: 1, 2: 1}),
    ]
    return out


def brute_force_cliques(n, edges):
    """All maximal cliques by subset enumeration (n <= 16)."""
    adj = [0] * n
    for a, b in edges:
        if a != b:
            adj[a] |= 1 << b
            adj[b] |= 1 << a
    used = [v for v in range(n) if adj[v]]
    cliques = []
    for mask in range(1, 1 << n):
        vs = [v for v in range(n) if mask >> v & 1]
        if any(not adj[v] for v in vs):
            continue
        if len(vs) < 2:
            continue
        ok = all((adj[v] & mask) == (mask & ~(1 << v)) for v in vs)
        if not ok:
            continue
        common = ~0
        for v in vs:
            common &= adj[v]
        if any(common >> u & 1 for u in used if not mask >> u & 1):
            continue
        cliques.append(tuple(vs))
    return cliques


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((n, n)) < p
    iu = np.triu_indices(n, 1)
    sel = m[iu]
    return list(zip(iu[0][sel].tolist(), iu[1][sel].tolist()))


def dense_core_graph(k, holes, n_sparse, p, seed):
    """K_k minus `holes` disjoint edges (2^holes maximal cliques of size
    k - holes ... ) joined to a sparse G(n_sparse, p) by random edges: the
    first dense vertex's edge roots mix roots with more than 64 candidates
    and roots with fewer, in one run sharing that vertex."""
    rng = np.random.default_rng(seed)
    cut = {(2 * i, 2 * i + 1) for i in range(holes)}
    edges = [(i, j) for i in range(k) for j in range(i + 1, k) if (i, j) not in cut]
    edges += [(k + a, k + b) for a, b in random_graph(n_sparse, p, seed + 1)]
    for v in range(k, k + n_sparse):
        for u in rng.choice(k, size=2, replace=False).tolist():
            edges.append((u, v))
    return edges


def hist_of(cliques):
    h = {}
    for c in cliques:
        h[len(c)] = h.get(len(c), 0) + 1
    return h


def expected_er_cliques(n, p):
    """Closed-form expectation of maximal cliques of size >= 2 in G(n, p)."""
    from rdmce_amd.workloads import expected_er_cliques as e
    return e(n, p)
