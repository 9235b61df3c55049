"""Host preprocessing (loader, orderings, CSR, .bin, generators) - CPU only.

Pinned by hand-derived expectations, by the CPU oracle's independent C
restatement, and - where oracle/_ref was built from the reference's own
sources - byte-for-byte against the reference's Graph::StoreIntoBin output
for every ordering."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import rdmce_amd as R
from graphs import random_graph
from oracle_lib import REF_MCE, OracleGraph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ORDERS = {"unchange": R.OrderType.UNCHANGED, "a": R.OrderType.ASC, "d": R.OrderType.DES, "deg": R.OrderType.DEG}


def load(path, order=None):
    g = R.Graph()
    g.LoadFromFile(path)
    if order is not None:
        g.SortByOrder(order)
    return g


def test_loader_edge_cases():
    """Comments, blank lines, self loops, duplicates, extra tokens, signs,
    overflow, a pair split over two lines, no trailing newline
    (graph.cc:83-104 with libstdc++ unsigned extraction)."""
    g = load(os.path.join(GOLDEN, "loader_cases.txt"))
    assert g.GetNumVertices() == 12 and g.GetNumEdges() == 8 and g.GetMaxDegree() == 2
    assert g.GetDegeneracy() == 2  # max degree until a DEG ordering
    assert list(g.GetLabels()) == [1, 2, 4, 5, 6, 7, 8, 9, 11, 12, 13, 14]
    adj = {0: [1], 1: [0, 4], 2: [5], 3: [5, 7], 4: [1], 5: [2, 3], 6: [7], 7: [3, 6], 8: [9], 9: [8], 10: [11],
           11: [10]}
    ro, co = g.GetRowOffset(), g.GetColIdx()
    for v, nb in adj.items():
        assert list(co[ro[v]:ro[v + 1]]) == nb


def test_loader_matches_oracle_restatement():
    for name in ["loader_cases.txt", "zachary.txt", "er2048_p0.008_s1.txt"]:
        path = os.path.join(GOLDEN, name)
        g = load(path)
        og = OracleGraph.load_coo(path)
        ro, co, lb = og.csr()
        assert np.array_equal(ro, g.GetRowOffset())
        assert np.array_equal(co, g.GetColIdx())
        assert np.array_equal(lb, g.GetLabels())


@pytest.mark.parametrize("name", ["zachary.txt", "er2048_p0.008_s1.txt", "loader_cases.txt"])
def test_degeneracy_order_matches_oracle_peel(name):
    path = os.path.join(GOLDEN, name)
    g = load(path, R.OrderType.DEG)
    og = OracleGraph.load_coo(path).order_degeneracy()
    ro, co, lb = og.csr()
    assert np.array_equal(ro, g.GetRowOffset()) and np.array_equal(co, g.GetColIdx())
    assert np.array_equal(lb, g.GetLabels())
    assert g.GetDegeneracy() == og.degeneracy


def test_degeneracy_properties_random():
    for seed in range(5):
        edges = random_graph(120, 0.15 + 0.1 * seed, seed)
        g = R.Graph.from_edges(edges)
        g.SortByOrder(R.OrderType.DEG)
        ro, co = g.GetRowOffset(), g.GetColIdx()
        d = g.GetDegeneracy()
        core = g.GetCoreNumbers()
        outdeg = [int(np.sum(co[ro[v]:ro[v + 1]] > v)) for v in range(g.GetNumVertices())]
        assert max(outdeg) <= d
        assert int(core.max()) == d
        assert np.all(np.diff(core.astype(np.int64)) >= 0)  # peel order is non-decreasing in core
        og = OracleGraph.from_edges(edges).order_degeneracy()
        assert np.array_equal(og.csr()[1], co)


def test_core_numbers_vs_networkx():
    nx = pytest.importorskip("networkx")
    for seed in range(4):
        edges = random_graph(200, 0.05 * (seed + 1), 100 + seed)
        g = R.Graph.from_edges(edges)
        core = dict(zip(g.GetLabels().tolist(), g.GetCoreNumbers().tolist()))
        G = nx.Graph(edges)
        assert core == nx.core_number(G)


def _ref_bin(path, order, tmp_path):
    out = tmp_path / "ref_out.txt"
    subprocess.run([REF_MCE, path, order, str(out)], check=True, capture_output=True)
    return (tmp_path / "ref_out.bin").read_bytes()


@pytest.mark.skipif(not os.path.exists(REF_MCE), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("name", ["zachary.txt", "er2048_p0.008_s1.txt", "loader_cases.txt"])
def test_bin_byte_identical_to_reference(name, order, tmp_path):
    """Loader + ordering + .bin writer against the reference's own code."""
    path = os.path.join(GOLDEN, name)
    ref = _ref_bin(path, order, tmp_path)
    g = load(path, ORDERS[order])
    mine = tmp_path / "mine.bin"
    g.StoreIntoBinExact(str(mine))
    assert mine.read_bytes() == ref


@pytest.mark.skipif(not os.path.exists(REF_MCE), reason="oracle/_ref not built (reference tree absent)")
def test_bin_random_graphs_vs_reference(tmp_path):
    rng = np.random.default_rng(7)
    for i in range(6):
        edges = random_graph(int(rng.integers(20, 300)), float(rng.choice([0.02, 0.1, 0.3])), 50 + i)
        path = tmp_path / f"g{i}.txt"
        path.write_text("".join(f"{a * 3 + 5} {b * 3 + 5}\n" for a, b in edges))
        for order in ORDERS:
            ref = _ref_bin(str(path), order, tmp_path)
            g = load(str(path), ORDERS[order])
            mine = tmp_path / "mine.bin"
            g.StoreIntoBinExact(str(mine))
            assert mine.read_bytes() == ref, (i, order)


def test_bin_roundtrip_and_name_rule(tmp_path):
    src = tmp_path / "graph.edges.txt"
    src.write_text(open(os.path.join(GOLDEN, "zachary.txt")).read())
    g = load(str(src), R.OrderType.DEG)
    out = g.StoreIntoBin(str(src))
    assert out == str(tmp_path / "graph.edges.bin")  # extension replaced (graph.cc:155-156)
    h = R.Graph()
    h.LoadFromBin(out)
    for a, b in [(g.GetRowOffset(), h.GetRowOffset()), (g.GetColIdx(), h.GetColIdx()),
                 (g.GetLabels(), h.GetLabels())]:
        assert np.array_equal(a, b)
    assert (h.GetNumVertices(), h.GetNumEdges(), h.GetMaxDegree(), h.GetDegeneracy()) == (34, 78, 17, 4)


def test_bin_truncated_is_an_error(tmp_path):
    g = load(os.path.join(GOLDEN, "zachary.txt"), R.OrderType.DEG)
    p = tmp_path / "z.bin"
    g.StoreIntoBinExact(str(p))
    data = p.read_bytes()
    (tmp_path / "t.bin").write_bytes(data[:len(data) // 2])
    h = R.Graph()
    with pytest.raises(R.MceError) as e:
        h.LoadFromBin(str(tmp_path / "t.bin"))
    assert e.value.status == 2


def test_missing_file_is_an_error():
    with pytest.raises(R.MceError) as e:
        R.Graph().LoadFromFile("/nonexistent/graph.txt")
    assert e.value.status == 2


def test_from_csr_validates():
    with pytest.raises(R.MceError):
        R.Graph.from_csr([0, 1, 1], [5])  # column out of range
    with pytest.raises(R.MceError):
        R.Graph.from_csr([0, 1], [0])  # self loop


def _digest(g):
    h = hashlib.sha256()
    for a in (g.GetRowOffset(), g.GetColIdx(), g.GetLabels()):
        h.update(a.tobytes())
    return h.hexdigest()


def test_generators_deterministic_and_thread_independent():
    code = ("import sys, hashlib; sys.path.insert(0, %r); import rdmce_amd as R;"
            "g = R.Graph.rmat(12, 16, seed=5); h = hashlib.sha256();"
            "[h.update(a.tobytes()) for a in (g.GetRowOffset(), g.GetColIdx(), g.GetLabels())];"
            "c = R.Graph.chung_lu_planted(20000, 2.1, 20.0, 20, 50, 200, 0.6, 4); h2 = hashlib.sha256();"
            "[h2.update(a.tobytes()) for a in (c.GetRowOffset(), c.GetColIdx())];"
            "print(h.hexdigest(), h2.hexdigest())" % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    outs = []
    for t in ("1", "3", "8"):
        env = dict(os.environ, RDMCE_THREADS=t)
        outs.append(subprocess.run(["python3", "-c", code], env=env, check=True, capture_output=True,
                                   text=True).stdout.strip())
    assert outs[0] == outs[1] == outs[2]


def test_er_generator_properties():
    g = R.Graph.erdos_renyi(2048, 0.008, 1)
    h = R.Graph()
    h.LoadFromFile(os.path.join(GOLDEN, "er2048_p0.008_s1.txt"))
    assert _digest(g) == _digest(h)  # the committed COO text is this generator's output
    m = g.GetNumEdges()
    exp = 0.008 * 2048 * 2047 / 2
    assert abs(m - exp) < 5 * np.sqrt(exp)


def test_rmat_generator_properties():
    g = R.Graph.rmat(14, 16, seed=3)
    assert g.GetNumVertices() <= 1 << 14
    lab = g.GetLabels()
    assert lab.max() < (1 << 14) and np.all(np.diff(lab.astype(np.int64)) > 0)
    deg = np.diff(g.GetRowOffset().astype(np.int64))
    # skew a > d: the lowest-id vertex is the hub
    assert deg[0] == deg.max()


def test_batch_params_in_range():
    for i in range(200):
        n, p, s = R.batch_graph_params(3, i)
        assert 32 <= n <= 1024 and 0.05 <= p < 0.5
    assert R.batch_graph_params(3, 7) == R.batch_graph_params(3, 7)
