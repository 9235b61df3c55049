"""GPU parity: the HIP path (through the C ABI) against the CPU oracle, the
committed goldens and closed forms.  Integer work: every comparison is
bit-exact (count, histogram, clique hash, clique list, per-root node counts).
"""
import os
import sys

import numpy as np
import pytest

import rdmce_amd as R
from graphs import closed_forms, expected_er_cliques, random_graph
from oracle_lib import OracleGraph, clique_hash

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def banner(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_for(g: R.Graph, root_mode=2, root_nodes=False, list_cap=0):
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    return og.mce(root_mode=root_mode, root_nodes=root_nodes, list_cap=list_cap)


def check_same(out, ref, *, lists=False, nodes=False):
    assert out.count == ref.count
    assert out.hist_dict() == ref.hist
    assert out.hash == ref.hash
    assert sum(out.hist_dict().values()) == out.count
    if lists:
        assert sorted(out.cliques) == sorted(tuple(sorted(c)) for c in ref.cliques)
        assert clique_hash(out.cliques) == out.hash
    if nodes:
        assert out.total_nodes == ref.total_nodes
        assert np.array_equal(out.root_nodes, ref.root_nodes)


def test_zachary_full():
    banner("zachary edge roots")
    g = R.Graph()
    g.LoadFromFile(os.path.join(GOLDEN, "zachary.txt"))
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, want_list=True, want_root_stats=True)
    ref = oracle_for(g, root_nodes=True, list_cap=1 << 16)
    assert out.count == 36
    assert out.hist_dict() == {2: 11, 3: 21, 4: 2, 5: 2}
    check_same(out, ref, lists=True, nodes=True)


@pytest.mark.parametrize("name,edges,hist", closed_forms(), ids=[c[0] for c in closed_forms()])
def test_closed_forms(name, edges, hist):
    banner(f"closed form {name}")
    g = R.Graph.from_edges(edges)
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, want_root_stats=True)
    assert out.hist_dict() == hist
    if sum(hist.values()) < 100000:
        check_same(out, oracle_for(g, root_nodes=True), nodes=True)


@pytest.mark.parametrize("n,p,seed", [(30, 0.3, 1), (60, 0.5, 2), (120, 0.7, 3), (300, 0.4, 4), (300, 0.05, 5),
                                      (80, 0.85, 6), (150, 0.6, 7), (400, 0.3, 8)])
def test_random_graphs(n, p, seed):
    banner(f"random n={n} p={p} seed={seed}")
    g = R.Graph.from_edges(random_graph(n, p, seed))
    g.SortByOrder(R.OrderType.DEG)
    lists = expected_er_cliques(n, p) < 3e5
    out = R.solve(g, want_root_stats=True, want_list=lists)
    ref = oracle_for(g, root_nodes=True, list_cap=(1 << 24) if lists else 0)
    check_same(out, ref, lists=lists, nodes=True)


def test_er_config1():
    banner("ER G(2048, 0.008) seed 1")
    g = R.Graph.erdos_renyi(2048, 0.008, 1)
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, want_list=True, want_root_stats=True)
    ref = oracle_for(g, root_nodes=True, list_cap=1 << 20)
    check_same(out, ref, lists=True, nodes=True)


@pytest.mark.parametrize("order", [R.OrderType.UNCHANGED, R.OrderType.ASC, R.OrderType.DES])
def test_order_invariance(order):
    banner(f"order invariance {int(order)}")
    edges = random_graph(150, 0.5, 11)
    g0 = R.Graph.from_edges(edges)
    g0.SortByOrder(R.OrderType.DEG)
    base = R.solve(g0)
    g = R.Graph.from_edges(edges)
    g.SortByOrder(order)
    out = R.solve(g)
    assert (out.count, out.hash, out.hist_dict()) == (base.count, base.hash, base.hist_dict())


def test_vertex_roots():
    banner("vertex roots")
    g = R.Graph.from_edges(random_graph(180, 0.6, 12))
    g.SortByOrder(R.OrderType.DEG)
    e = R.solve(g)
    v = R.solve(g, root_mode=R.ROOT_VERTEX, want_root_stats=True)
    ref = oracle_for(g, root_mode=1, root_nodes=True)
    check_same(v, ref, nodes=True)
    assert (v.count, v.hash, v.hist_dict()) == (e.count, e.hash, e.hist_dict())


def test_deferred_pass():
    banner("deferred second pass")
    g = R.Graph.from_edges(random_graph(250, 0.45, 13))
    g.SortByOrder(R.OrderType.DEG)
    base = R.solve(g, want_root_stats=True)
    small = R.solve(g, slot_bytes=4096, want_root_stats=True)
    assert small.passes == 2 and small.deferred_roots > 0
    assert (small.count, small.hash, small.hist_dict()) == (base.count, base.hash, base.hist_dict())
    assert np.array_equal(small.root_nodes, base.root_nodes)


def test_custom_schedule():
    banner("custom root permutation")
    g = R.Graph.from_edges(random_graph(160, 0.55, 14))
    g.SortByOrder(R.OrderType.DEG)
    base = R.solve(g, want_root_stats=True)
    perm = np.random.default_rng(0).permutation(g.GetNumEdges())
    out = R.solve(g, root_order=perm, want_root_stats=True)
    assert (out.count, out.hash, out.hist_dict()) == (base.count, base.hash, base.hist_dict())
    assert np.array_equal(out.root_nodes, base.root_nodes)


def test_plan_rerun():
    banner("plan rerun")
    g = R.Graph.from_edges(random_graph(140, 0.6, 15))
    g.SortByOrder(R.OrderType.DEG)
    plan = R.Plan(g)
    a = plan.run()
    b = plan.run()
    plan.free()
    assert (a.count, a.hash, a.hist_dict()) == (b.count, b.hash, b.hist_dict())


@pytest.mark.parametrize("tau", [1, 2, 4, 24, -1])
def test_sharing_tree_invariance(tau):
    """Root-down exposure must not change the tree: counts, hash, histogram
    and per-root node counts identical for every sharing bound."""
    banner(f"sharing tau={tau}")
    g = R.Graph.from_edges(random_graph(300, 0.5, 21))
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, share_bound=tau, want_root_stats=True)
    ref = oracle_for(g, root_nodes=True)
    check_same(out, ref, nodes=True)
    if tau in (1, 2):
        assert out.shared_tasks > 0


def test_sharing_single_heavy_root():
    """One root holds all the work (Moon-Moser K_{3x9} plus a pendant):
    other waves can only help by stealing."""
    banner("sharing heavy root")
    from graphs import moon_moser
    g = R.Graph.from_edges(moon_moser(9))
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, share_bound=2, want_root_stats=True)
    assert out.hist_dict() == {9: 3 ** 9}
    ref = oracle_for(g, root_nodes=True)
    check_same(out, ref, nodes=True)
    assert out.shared_tasks > 0


@pytest.mark.parametrize("tau", [-1, 2])
def test_register_paths_match_bitmap_tiers(tau):
    """Vertex roots of G(300, 0.3) have 65..81 candidates (two-word register
    path) or fewer (single-word); the same graph with the register paths off
    runs every root in the bitmap tiers.  Both equal the oracle, per-root
    node counts included."""
    banner(f"register paths vs bitmap tiers tau={tau}")
    g = R.Graph.from_edges(random_graph(300, 0.3, 32))
    g.SortByOrder(R.OrderType.DEG)
    ref = oracle_for(g, root_mode=1, root_nodes=True)
    for wp in (True, False):
        out = R.solve(g, root_mode=R.ROOT_VERTEX, share_bound=tau, word_path=wp, want_root_stats=True)
        check_same(out, ref, nodes=True)


def test_rmat20_core499_full_size_golden():
    """The bench workload at full size against the committed CPU-oracle
    golden (tests/golden/make_workload_golden.py): 13.2M roots, ~1.0e9 cliques."""
    import json
    banner("rmat20 core<=499 full size")
    with open(os.path.join(GOLDEN, "rmat20_s2_core499.json")) as f:
        gd = json.load(f)
    g = R.Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=2)
    g.SortByOrder(R.OrderType.DEG)
    core = g.GetCoreNumbers()
    roots = g.roots(R.ROOT_EDGE)
    v0 = (roots >> np.uint64(32)).astype(np.int64)
    end = int(np.searchsorted(core[v0], gd["core_max"], side="right"))
    assert end == gd["roots"]
    out = R.solve(g, root_order=np.arange(end, dtype=np.uint64)[::-1].copy())
    assert out.count == gd["count"]
    assert out.hash == gd["hash"]
    assert out.hist_dict() == {int(k): v for k, v in gd["hist"].items()}
    assert out.total_nodes == gd["tree_nodes"]


def test_cl4m_core300_full_size_golden():
    """Chung-Lu power-law graph with planted communities (n = 4e6), every
    edge root with core(v0) <= 300 (48.2M roots, 2.19e9 cliques), against
    the committed CPU-oracle golden (tests/golden/make_workload_golden.py)."""
    import json
    from rdmce_amd import workloads
    banner("cl4m core<=300 full size")
    with open(os.path.join(GOLDEN, "cl4m_s4_core300.json")) as f:
        gd = json.load(f)
    w = workloads.load("cl4m-core300")
    assert w.n_roots == gd["roots"]
    out = R.solve(w.graph, root_order=w.plain_order())
    assert out.count == gd["count"]
    assert out.hash == gd["hash"]
    assert out.hist_dict() == {int(k): v for k, v in gd["hist"].items()}
    assert out.total_nodes == gd["tree_nodes"]


def test_rmat24_core500_full_size_golden():
    """R-MAT scale 24 (n = 8.9e6 after dropping isolated vertices, 2.6e8
    edges), every edge root with core(v0) <= 500 (2.1e8 roots, 1.02e10
    cliques, 2.3e10 search-tree nodes), against the committed CPU-oracle
    golden (tests/golden/make_workload_golden.py; half an hour on 8 host
    threads)."""
    import json
    from rdmce_amd import workloads
    banner("rmat24 core<=500 full size")
    with open(os.path.join(GOLDEN, "rmat24_s5_core500.json")) as f:
        gd = json.load(f)
    w = workloads.load("rmat24-core500")
    assert (w.graph.GetNumVertices(), w.graph.GetNumEdges(), w.n_roots) == (gd["n"], gd["m"], gd["roots"])
    banner(f"rmat24 graph built: {w.n_roots} roots")
    out = R.solve(w.graph, root_order=w.plain_order())
    assert out.count == gd["count"]
    assert out.hash == gd["hash"]
    assert out.hist_dict() == {int(k): v for k, v in gd["hist"].items()}
    assert out.total_nodes == gd["tree_nodes"]


def test_batch_per_graph_results():
    """Many graphs in one launch: per-graph count, hash and histogram equal
    each graph's own oracle result."""
    banner("batch per-graph")
    graphs, refs = [], []
    for i in range(40):
        n, p, s = R.batch_graph_params(3, i, 32, 160, 0.05, 0.5)
        g = R.Graph.erdos_renyi(n, p, s)
        g.SortByOrder(R.OrderType.DEG)
        graphs.append(g)
        refs.append(oracle_for(g))
    b = R.Batch(graphs)
    out = b.run()
    b.free()
    for i, ref in enumerate(refs):
        assert int(out.count[i]) == ref.count, i
        assert int(out.hash[i]) == ref.hash, i
        assert out.hist_dict(i) == ref.hist, i
    assert out.total_count == sum(r.count for r in refs)


def test_batch_graph_replay_matches_eager():
    """Captured hipGraph replays (reset fills + kernel) equal the eager run,
    with eager work and synchronisation between replays."""
    banner("batch graph replay")
    graphs = []
    for i in range(25):
        n, p, s = R.batch_graph_params(3, 100 + i, 32, 200, 0.05, 0.5)
        g = R.Graph.erdos_renyi(n, p, s)
        g.SortByOrder(R.OrderType.DEG)
        graphs.append(g)
    eager = R.Batch(graphs).run()
    b = R.Batch(graphs, use_graph=True)
    for rep in range(4):
        out = b.run()
        assert out.graph_replay
        assert np.array_equal(out.count, eager.count) and np.array_equal(out.hash, eager.hash)
        assert np.array_equal(out.hist, eager.hist)
        R.solve(graphs[rep])  # unrelated eager work between replays
    b.free()


def test_plan_graph_replay_and_eager_split():
    banner("plan graph replay + eager split")
    g = R.Graph.from_edges(random_graph(260, 0.5, 31))
    g.SortByOrder(R.OrderType.DEG)
    base = R.solve(g, want_root_stats=True)
    m = g.GetNumEdges()
    order = np.arange(m, dtype=np.uint64)[::-1].copy()
    flags = np.zeros(m, dtype=np.uint8)
    flags[: m // 10] = 1
    plan = R.Plan(g, root_order=order, eager=flags, use_graph=True, want_root_stats=True)
    for _ in range(3):
        out = plan.run()
        assert (out.count, out.hash, out.hist_dict()) == (base.count, base.hash, base.hist_dict())
        assert np.array_equal(out.root_nodes, base.root_nodes)
    plan.free()


def test_autotune_keeps_results():
    banner("autotune")
    g = R.Graph.from_edges(random_graph(300, 0.45, 32))
    g.SortByOrder(R.OrderType.DEG)
    base = R.solve(g)
    plan = R.Plan(g)
    t = plan.autotune(stride=2)
    assert t["trials"] >= 10 and t["best_ms"] <= t["base_ms"] * 1.0001
    out = plan.run()
    plan.free()
    assert (out.count, out.hash, out.hist_dict()) == (base.count, base.hash, base.hist_dict())
