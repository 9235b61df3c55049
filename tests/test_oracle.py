"""Pins the CPU oracle before it is trusted as the checker (CPU only):
networkx goldens, brute force over all subsets, closed forms, the textbook
list-based Tomita in the same library, the reference's own vertex-rooted
Tomita solver built from its sources (oracle/_ref), and invariants.

A plausible oracle mistake - a wrong pivot, a dropped X row, a missing
aggressive-generation guard, a wrong hash term - changes at least one of
these numbers; tests/test_oracle.py::test_planted_mistakes_are_caught checks
that by planting such mistakes in a Python copy of the canonical tree."""
import json
import os
import subprocess

import numpy as np
import pytest

from graphs import brute_force_cliques, closed_forms, expected_er_cliques, hist_of, random_graph
from oracle_lib import REF_MCE, OracleGraph, clique_hash, splitmix64

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def as_hist(h):
    return {int(k): int(v) for k, v in h.items()}


def test_splitmix64_known_answers():
    # SplitMix64 reference outputs for state 0 (first three outputs of the
    # generator seeded with 0, from Vigna's published splitmix64.c)
    assert splitmix64(0) == 0xE220A8397B1DCDAF
    assert splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    assert splitmix64(2 * 0x9E3779B97F4A7C15 % (1 << 64)) == 0x06C45D188009454F


@pytest.mark.parametrize("mode", [2, 1])
def test_zachary_golden(mode):
    gd = load_golden("zachary.json")
    og = OracleGraph.load_coo(os.path.join(GOLDEN, "zachary.txt")).order_degeneracy()
    r = og.mce(root_mode=mode, list_cap=1 << 16)
    assert r.count == gd["count"] == 36
    assert r.hist == as_hist(gd["hist"]) == {2: 11, 3: 21, 4: 2, 5: 2}
    assert r.hash == gd["hash"] == clique_hash(r.cliques)


def test_er_config1_golden():
    gd = load_golden("er2048_p0.008_s1.json")
    og = OracleGraph.load_coo(os.path.join(GOLDEN, "er2048_p0.008_s1.txt")).order_degeneracy()
    r = og.mce(list_cap=1 << 20)
    assert (r.count, r.hist, r.hash) == (gd["count"], as_hist(gd["hist"]), gd["hash"])
    import hashlib
    cl = sorted(tuple(sorted(c)) for c in r.cliques)
    assert hashlib.sha256(json.dumps(cl).encode()).hexdigest() == gd["sha256_sorted"]


def test_random_small_goldens():
    cases = load_golden("random_small.json")
    assert len(cases) > 50
    for c in cases:
        e = np.array(c["edges"], dtype=np.int64)
        og = OracleGraph.from_pairs(e[:, 0], e[:, 1]).order_degeneracy()
        for mode in (2, 1):
            r = og.mce(root_mode=mode)
            assert (r.count, r.hist, r.hash) == (c["count"], as_hist(c["hist"]), c["hash"]), (c["n"], c["p"], mode)


@pytest.mark.parametrize("seed", range(12))
def test_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(4, 15))
    edges = random_graph(n, float(rng.choice([0.2, 0.4, 0.6, 0.8])), seed)
    if not edges:
        return
    bf = brute_force_cliques(n, edges)
    og = OracleGraph.from_edges(edges).order_degeneracy()
    r = og.mce(list_cap=1 << 16)
    assert sorted(tuple(sorted(c)) for c in r.cliques) == sorted(bf)
    assert r.hist == hist_of(bf) and r.hash == clique_hash(bf)


@pytest.mark.parametrize("name,edges,hist", closed_forms(), ids=[c[0] for c in closed_forms()])
def test_closed_forms(name, edges, hist):
    og = OracleGraph.from_edges(edges).order_degeneracy()
    assert og.mce().hist == hist
    if sum(hist.values()) < 50000:
        assert og.mce(root_mode=1).hist == hist


@pytest.mark.parametrize("n,p,seed", [(40, 0.3, 1), (70, 0.5, 2), (60, 0.8, 3), (150, 0.1, 4)])
def test_tomita_crosscheck(n, p, seed):
    og = OracleGraph.from_edges(random_graph(n, p, seed)).order_degeneracy()
    a, b = og.mce(), og.tomita()
    assert (a.count, a.hist, a.hash) == (b.count, b.hist, b.hash)


def test_order_invariance():
    edges = random_graph(90, 0.4, 9)
    base = OracleGraph.from_edges(edges).order_degeneracy().mce()
    unordered = OracleGraph.from_edges(edges).mce()  # ascending-label order
    assert (base.count, base.hist, base.hash) == (unordered.count, unordered.hist, unordered.hash)


def test_root_ranges_are_additive():
    og = OracleGraph.from_edges(random_graph(100, 0.4, 10)).order_degeneracy()
    full = og.mce(root_nodes=True)
    m = og.num_roots()
    parts = [og.mce(root_begin=b, root_end=min(m, b + 97), root_nodes=True) for b in range(0, m, 97)]
    assert sum(p.count for p in parts) == full.count
    assert sum(p.hash for p in parts) % (1 << 64) == full.hash
    assert np.array_equal(np.concatenate([p.root_nodes for p in parts]), full.root_nodes)
    ids = np.random.default_rng(0).permutation(m)
    a, b = og.mce_roots(ids[: m // 3]), og.mce_roots(ids[m // 3:])
    assert a.count + b.count == full.count and (a.hash + b.hash) % (1 << 64) == full.hash
    assert {k: a.hist.get(k, 0) + b.hist.get(k, 0) for k in full.hist} == full.hist


def test_thread_count_invariance():
    og = OracleGraph.from_edges(random_graph(150, 0.5, 11)).order_degeneracy()
    a = og.mce(threads=1, root_nodes=True)
    b = og.mce(threads=7, root_nodes=True)
    assert (a.count, a.hash, a.hist, a.total_nodes) == (b.count, b.hash, b.hist, b.total_nodes)
    assert np.array_equal(a.root_nodes, b.root_nodes)


def test_er_expectation_closed_form():
    """Mean clique count over G(n,p) samples matches the closed form."""
    n, p = 60, 0.3
    counts = [OracleGraph.from_edges(random_graph(n, p, 1000 + s)).order_degeneracy().mce().count
              for s in range(40)]
    exp = expected_er_cliques(n, p)
    assert abs(np.mean(counts) - exp) < 4 * np.std(counts) / np.sqrt(len(counts)) + 0.02 * exp


@pytest.mark.skipif(not os.path.exists(REF_MCE), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("name", ["zachary.txt", "er2048_p0.008_s1.txt"])
def test_reference_solver_count(name, tmp_path):
    """Clique count of the reference's own CPU solver (RDMCE/src/mce.cc:64-197)."""
    path = os.path.join(GOLDEN, name)
    out = subprocess.run([REF_MCE, path, "deg", str(tmp_path / "x.txt"), "count"], check=True,
                         capture_output=True, text=True).stdout
    ref_count = int([ln for ln in out.splitlines() if ln.startswith("count ")][0].split()[1])
    og = OracleGraph.load_coo(path).order_degeneracy()
    assert og.mce().count == ref_count


@pytest.mark.skipif(not os.path.exists(REF_MCE), reason="oracle/_ref not built (reference tree absent)")
def test_reference_solver_random_counts(tmp_path):
    for i, (n, p) in enumerate([(50, 0.5), (120, 0.2), (80, 0.7), (300, 0.05)]):
        edges = random_graph(n, p, 300 + i)
        path = tmp_path / f"r{i}.txt"
        path.write_text("".join(f"{a} {b}\n" for a, b in edges))
        out = subprocess.run([REF_MCE, str(path), "deg", str(tmp_path / "x.txt"), "count"], check=True,
                             capture_output=True, text=True).stdout
        ref_count = int([ln for ln in out.splitlines() if ln.startswith("count ")][0].split()[1])
        assert OracleGraph.load_coo(str(path)).order_degeneracy().mce().count == ref_count


# --- planted mistakes --------------------------------------------------------
def _py_canonical(n, edges, mistake=None, labels=None):
    """Python copy of the canonical edge-rooted tree (small graphs), with an
    optional planted mistake, returning (hist, hash, nodes)."""
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    # the caller passes a graph already in degeneracy order: position = id
    order = [v for v in range(n) if adj[v]]
    pos = {v: v for v in order}
    cl, nodes = [], 0
    for v0 in order:
        for v1 in sorted(adj[v0], key=pos.get):
            if pos[v1] <= pos[v0]:
                continue
            C = adj[v0] & adj[v1]
            P0 = sorted((c for c in C if pos[c] > pos[v1]), key=pos.get)
            X0 = sorted((c for c in C if pos[c] < pos[v1]), key=pos.get)
            if not P0:
                if not X0:
                    cl.append((v0, v1))
                continue
            Xq = [x for x in X0 if adj[x] & set(P0)]
            if mistake == "drop_x":
                Xq = Xq[1:]
            rows = P0 + Xq

            def node(R, P):
                nonlocal nodes
                nodes += 1
                cand = [u for i, u in enumerate(rows) if all(r in adj[u] for r in R)]
                if not P:
                    if not cand or mistake == "no_leaf_check":
                        cl.append(tuple([v0, v1] + list(R)))
                    return
                key = (lambda u: (-len([q for q in P if q not in adj[u]]), rows.index(u))) if mistake == "pivot_max" \
                    else (lambda u: (len([q for q in P if q not in adj[u]]), rows.index(u)))
                piv = min(cand, key=key)
                if len([q for q in P if q not in adj[piv]]) == 0:
                    return
                if mistake == "pivot_max" and min(len([q for q in P if q not in adj[u]]) for u in cand) == 0:
                    return
                V = [q for q in P if q not in adj[piv]]
                Pc = list(P)
                for c in V:
                    Pc.remove(c)
                    R2 = list(R) + [c]
                    P2 = [q for q in Pc if q in adj[c]]
                    if len(P2) > 1 and mistake != "no_aggressive":
                        M = [i for i in P2 if all(j in adj[i] for j in P2 if j != i)]
                        R2 += M
                        P2 = [q for q in P2 if q not in M]
                    node(R2, P2)
            node([], P0)
    if labels is not None:
        cl = [tuple(int(labels[v]) for v in c) for c in cl]
    return hist_of(cl), clique_hash(cl), nodes


def test_planted_mistakes_are_caught():
    """The checks above catch plausible mistakes: the Python copy of the
    canonical tree agrees with the C oracle, and each planted mistake
    changes the node count, the histogram or the hash."""
    og = OracleGraph.from_edges(random_graph(40, 0.5, 77)).order_degeneracy()
    ro, co, lb = og.csr()
    n = og.n
    edges = [(v, int(u)) for v in range(n) for u in co[ro[v]:ro[v + 1]] if u > v]
    r = og.mce()
    h, hs, nodes = _py_canonical(n, edges, labels=lb)
    assert (h, hs) == (r.hist, r.hash)
    assert nodes == r.total_nodes  # the same canonical tree, node for node
    for mistake in ["drop_x", "no_aggressive", "no_leaf_check", "pivot_max"]:
        h2, hs2, nodes2 = _py_canonical(n, edges, mistake, labels=lb)
        assert (h2, hs2, nodes2) != (h, hs, nodes), mistake


@pytest.mark.parametrize("n,p,seed", [(60, 0.3, 1), (90, 0.5, 2), (200, 0.05, 3)])
def test_build_terms_brute_force(n, p, seed):
    """The algorithmic build terms (d(v0) + d(v1), degrees of P0's members,
    rows x words) against a set-based recount of every edge root, and the
    root ranges add up."""
    from rdmce_amd.graph import Graph, OrderType
    g = Graph.from_edges(random_graph(n, p, seed))
    g.SortByOrder(OrderType.DEG)
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    ro, col, nv = g.GetRowOffset(), g.GetColIdx(), g.GetNumVertices()
    N = [set(col[ro[v]:ro[v + 1]].tolist()) for v in range(nv)]
    deg = pdeg = rw = roots = 0
    for v0 in range(nv):
        for v1 in sorted(x for x in N[v0] if x > v0):
            c = N[v0] & N[v1]
            P = {x for x in c if x > v1}
            xq = sum(1 for x in c if x < v1 and N[x] & P) if P else 0
            deg += len(N[v0]) + len(N[v1])
            pdeg += sum(len(N[u]) for u in P)
            rw += (len(P) + xq) * ((len(P) + 63) // 64)
            roots += 1
    t = og.build_terms(threads=3)
    assert (t["roots"], t["deg"], t["pdeg"], t["rows_w"]) == (roots, deg, pdeg, rw)
    assert t["build_bytes"] == 4 * deg + 4 * pdeg + 8 * rw
    m = roots
    a, b = og.build_terms(0, m // 3), og.build_terms(m // 3, m)
    assert all(a[k] + b[k] == t[k] for k in ("roots", "deg", "pdeg", "rows_w"))
