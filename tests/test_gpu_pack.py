"""GPU parity of the packed search (pack_kernel.hip: many small roots per
wavefront, one search-tree node per lane) against the CPU oracle and the
one-root-per-wavefront kernel.  Integer results: bit-exact count,
histogram, clique hash, clique list and per-root node counts; the bitmap
words the search reads (search_bytes) equal the other kernel's, node for
node."""
import json
import os
import sys

import numpy as np
import pytest

import rdmce_amd as R
from rdmce_amd._lib import MceError
from graphs import closed_forms, dense_core_graph, expected_er_cliques, random_graph
from oracle_lib import OracleGraph, clique_hash

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def banner(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_for(g, root_mode=2, root_nodes=True, list_cap=0):
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    return og.mce(root_mode=root_mode, root_nodes=root_nodes, list_cap=list_cap)


def check_same(out, ref, lists=False):
    assert (out.count, out.hash, out.hist_dict()) == (ref.count, ref.hash, ref.hist)
    assert out.total_nodes == ref.total_nodes
    assert np.array_equal(out.root_nodes, ref.root_nodes)
    if lists:
        assert sorted(out.cliques) == sorted(tuple(sorted(c)) for c in ref.cliques)
        assert clique_hash(out.cliques) == out.hash


def test_zachary_packed():
    banner("pack zachary")
    g = R.Graph()
    g.LoadFromFile(os.path.join(GOLDEN, "zachary.txt"))
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, pack=True, want_list=True, want_root_stats=True)
    assert out.count == 36 and out.hist_dict() == {2: 11, 3: 21, 4: 2, 5: 2}
    check_same(out, oracle_for(g, list_cap=1 << 16), lists=True)


@pytest.mark.parametrize("name,edges,hist", closed_forms(), ids=[c[0] for c in closed_forms()])
def test_closed_forms_packed(name, edges, hist):
    banner(f"pack closed form {name}")
    g = R.Graph.from_edges(edges)
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, pack=True, want_root_stats=True)
    assert out.hist_dict() == hist
    if sum(hist.values()) < 100000:
        check_same(out, oracle_for(g))


@pytest.mark.parametrize("n,p,seed,slots", [(30, 0.3, 1, 4), (60, 0.5, 2, 1), (120, 0.7, 3, 4), (300, 0.4, 4, 8),
                                            (300, 0.05, 5, 4), (80, 0.85, 6, 2), (150, 0.6, 7, 4),
                                            (400, 0.3, 8, 4)])
def test_random_graphs_packed(n, p, seed, slots):
    """Roots of one, two and more words (the 400-vertex graph has roots
    wider than 128 candidates, which reach the one-root-per-wave kernel)."""
    banner(f"pack random n={n} p={p} seed={seed} slots={slots}")
    g = R.Graph.from_edges(random_graph(n, p, seed))
    g.SortByOrder(R.OrderType.DEG)
    lists = expected_er_cliques(n, p) < 3e5
    out = R.solve(g, pack=True, pack_slots=slots, want_root_stats=True, want_list=lists)
    check_same(out, oracle_for(g, list_cap=(1 << 24) if lists else 0), lists=lists)
    plain = R.solve(g, want_root_stats=True)
    assert out.search_bytes == plain.search_bytes


@pytest.mark.parametrize("k,slots", [(90, 2), (150, 4)])
def test_wide_runs_packed(k, slots):
    """A dense vertex whose run of edge roots mixes one-word, two-word and
    (k = 150) wider roots: the second launch takes the two-word ones as runs
    derived from one vertex bitmap, the rest reach the one-root-per-wave
    kernel."""
    banner(f"pack wide runs K{k}")
    g = R.Graph.from_edges(dense_core_graph(k, 5, 400, 0.03, 11))
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, pack=True, pack_slots=slots, want_root_stats=True, want_list=True)
    check_same(out, oracle_for(g, list_cap=1 << 20), lists=True)
    plain = R.solve(g, want_root_stats=True)
    assert out.search_bytes == plain.search_bytes


@pytest.mark.parametrize("pack", [False, True])
def test_choose_root_mode(pack):
    """The granularity chooser times vertex roots against the edge roots of
    the same vertices; both enumerate the same cliques, which it checks."""
    banner(f"choose_root_mode pack={pack}")
    g = R.Graph.from_edges(random_graph(400, 0.1, 21))
    g.SortByOrder(R.OrderType.DEG)
    ch = R.choose_root_mode(g, stride=3, pack=pack)
    assert ch["root_mode"] in (R.ROOT_VERTEX, R.ROOT_EDGE)
    assert ch["vertex_ms"] > 0 and ch["edge_ms"] > 0 and ch["sample_vertices"] == 134
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    ref = og.mce_roots(np.arange(0, 400, 3, dtype=np.uint64), root_mode=1)
    assert ch["sample_cliques"] == ref.count


EDGE_CASES = {"one edge": [(0, 1)], "matching": [(2 * i, 2 * i + 1) for i in range(40)],
              "star": [(0, i) for i in range(1, 70)], "path": [(i, i + 1) for i in range(100)],
              "two triangles on an edge": [(0, 1), (1, 2), (0, 2), (1, 3), (0, 3)]}


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
def test_edge_case_graphs_packed(name):
    """Graphs whose roots are all 2-cliques, single triangles or empty P:
    the packed search settles them during the build, as the oracle does."""
    banner(f"pack edge case {name}")
    g = R.Graph.from_edges(EDGE_CASES[name])
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, pack=True, want_root_stats=True, want_list=True)
    check_same(out, oracle_for(g, list_cap=1 << 12), lists=True)


def test_empty_graph_packed():
    banner("pack empty graph")
    g = R.Graph.from_edges([])
    assert g.GetNumVertices() == 0
    out = R.solve(g, pack=True)
    assert (out.count, out.hist_dict()) == (0, {})


def test_er_config1_packed():
    banner("pack ER G(2048, 0.008) seed 1")
    g = R.Graph.erdos_renyi(2048, 0.008, 1)
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, pack=True, want_list=True, want_root_stats=True)
    check_same(out, oracle_for(g, list_cap=1 << 20), lists=True)


def test_vertex_roots_packed():
    banner("pack vertex roots")
    g = R.Graph.from_edges(random_graph(300, 0.3, 32))
    g.SortByOrder(R.OrderType.DEG)
    out = R.solve(g, pack=True, root_mode=R.ROOT_VERTEX, want_root_stats=True)
    check_same(out, oracle_for(g, root_mode=1))


def test_batch_packed_and_replayed():
    """Per-graph results of a packed batch, eager and as replayed hipGraphs
    (with unrelated eager work between replays)."""
    banner("pack batch + replay")
    graphs, refs = [], []
    for i in range(40):
        n, p, s = R.batch_graph_params(3, i, 32, 200, 0.05, 0.5)
        g = R.Graph.erdos_renyi(n, p, s)
        g.SortByOrder(R.OrderType.DEG)
        graphs.append(g)
        refs.append(oracle_for(g, root_nodes=False))
    b = R.Batch(graphs, pack=True, use_graph=True)
    for rep in range(3):
        out = b.run()
        for i, ref in enumerate(refs):
            assert int(out.count[i]) == ref.count, (rep, i)
            assert int(out.hash[i]) == ref.hash, (rep, i)
            assert out.hist_dict(i) == ref.hist, (rep, i)
        R.solve(graphs[rep])
    b.free()


def test_plan_rerun_packed():
    banner("pack plan rerun")
    g = R.Graph.from_edges(random_graph(260, 0.5, 31))
    g.SortByOrder(R.OrderType.DEG)
    ref = oracle_for(g)
    plan = R.Plan(g, pack=True, want_root_stats=True, use_graph=True)
    for _ in range(3):
        check_same(plan.run(), ref)
    plan.free()


@pytest.mark.parametrize("name,golden", [("rmat20-core499", "rmat20_s2_core499.json"),
                                         ("cl4m-core300", "cl4m_s4_core300.json"),
                                         ("rmat24-core500", "rmat24_s5_core500.json")])
def test_full_size_golden_packed(name, golden):
    from rdmce_amd import workloads
    with open(os.path.join(GOLDEN, golden)) as f:
        gd = json.load(f)
    w = workloads.load(name)
    assert w.n_roots == gd["roots"]
    banner(f"pack {name}: {w.n_roots} roots")
    out = R.solve(w.graph, root_order=w.plain_order(), pack=True)
    assert out.count == gd["count"]
    assert out.hash == gd["hash"]
    assert out.hist_dict() == {int(k): v for k, v in gd["hist"].items()}
    assert out.total_nodes == gd["tree_nodes"]
    banner(f"pack {name}: kernel {out.kernel_ms:.1f} ms")


def test_autotune_packed_keeps_results():
    banner("pack autotune")
    g = R.Graph.from_edges(random_graph(300, 0.45, 32))
    g.SortByOrder(R.OrderType.DEG)
    ref = oracle_for(g)
    plan = R.Plan(g, pack=True, want_root_stats=True)
    t = plan.autotune(stride=2)
    assert t["trials"] >= 8 and t["best_ms"] <= t["base_ms"] * 1.0001
    assert t["pack_slots"] in (1, 2, 4, 8) and t["pack_waves_per_cu"] in (12, 16, 20, 24)
    check_same(plan.run(), ref)
    plan.free()


def test_learned_group_schedule_packed():
    """Learned run order, run splits and routed roots on a graph whose roots
    vary widely in cost: same results and per-root node counts."""
    from rdmce_amd import learned
    banner("pack learned group schedule")
    g = R.Graph.rmat(14, 16, 0.57, 0.19, 0.19, seed=7)
    g.SortByOrder(R.OrderType.DEG)
    m = g.GetNumEdges()
    ref = oracle_for(g)
    pred, model, rep = learned.fit_and_predict(g, m, stride=4)
    model.free()
    v0 = (g.roots(R.ROOT_EDGE) >> np.uint64(32)).astype(np.int64)
    sc = learned.learned_group_schedule(pred, np.arange(m, dtype=np.uint64), v0, 64, chunk_share=0.05,
                                        route_share=0.2)
    assert sc.eager.sum() > 0 and sc.runs.size - 1 > np.unique(v0).size  # some routed, some runs split
    out = R.solve(g, root_order=sc.order, eager=sc.eager, runs=sc.runs, pack=True, want_root_stats=True)
    check_same(out, ref)


def test_runs_must_share_a_first_vertex():
    banner("pack runs validation")
    g = R.Graph.from_edges(random_graph(60, 0.5, 2))
    g.SortByOrder(R.OrderType.DEG)
    m = g.GetNumEdges()
    order = np.arange(m, dtype=np.uint64)[::-1].copy()
    with pytest.raises(MceError, match="mixes first vertices"):
        R.Plan(g, root_order=order, runs=np.array([0, m], dtype=np.uint64), pack=True)
    with pytest.raises(MceError, match="start at 0"):
        R.Plan(g, root_order=order, runs=np.array([1, m], dtype=np.uint64), pack=True)
