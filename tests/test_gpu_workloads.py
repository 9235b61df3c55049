"""GPU parity of the full-size workloads through the batched, learned and
autotuned paths, against committed CPU-oracle goldens (tests/golden/
make_workload_golden.py, make_batch_sample_golden.py).  Integer results:
bit-exact.

  * the 10,000-graph G(n, p) batch family (seed 3): all 6,315 graphs with
    expected count <= 1e6 in one persistent launch replayed as a captured
    hipGraph, and a stratified sample of the heavier graphs in one launch;
  * Chung-Lu + planted communities (n = 4e6, seed 4) and R-MAT scale 24
    (seed 5), core-bounded: the learned schedule (GPU-trained cost model)
    with the autotuned settings.
"""
import json
import os
import sys

import numpy as np
import pytest

import rdmce_amd as R
from rdmce_amd import learned, workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def banner(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_batch10k_e1e6_one_replayed_launch():
    """All 6,315 graphs of the family with expected count <= 1e6, one
    persistent launch per run, captured once and replayed: per-graph
    (count, hash) digest, total and summed histogram equal the oracle's."""
    gd = load_golden("batch10k_s3_e1e6.json")
    bw = workloads.load_batch("batch10k-e1e6")
    banner(f"batch10k-e1e6: {len(bw.graphs)} graphs")
    assert len(bw.graphs) == gd["graphs"]
    b = R.Batch(bw.graphs, use_graph=True)
    try:
        for rep in range(2):
            out = b.run()
            assert out.graph_replay
            assert out.total_count == gd["total_count"], rep
            assert workloads.batch_digest(out.count, out.hash) == gd["digest"], rep
            hist = {int(k): int(v) for k, v in enumerate(out.hist.sum(axis=0)) if v}
            assert hist == {int(k): v for k, v in gd["hist"].items()}, rep
            assert out.total_nodes == gd["tree_nodes"], rep
    finally:
        b.free()


def test_batch10k_heavy_sample_one_launch():
    """A stratified sample of the heavier graphs of the family (expected
    counts 1e6 .. 1e10) in one launch: per-graph count, hash and histogram
    equal the oracle's."""
    gd = load_golden("batch10k_s3_heavy_sample.json")
    graphs = []
    for rec in gd["graphs"]:
        n, p, s = R.batch_graph_params(3, rec["index"], 32, 1024, 0.05, 0.5)
        assert (n, p) == (rec["n"], rec["p"])
        g = R.Graph.erdos_renyi(n, p, s)
        g.SortByOrder(R.OrderType.DEG)
        graphs.append(g)
    banner(f"batch10k heavy sample: {len(graphs)} graphs, {gd['total_count']} cliques")
    b = R.Batch(graphs)
    try:
        out = b.run()
    finally:
        b.free()
    for i, rec in enumerate(gd["graphs"]):
        assert int(out.count[i]) == rec["count"], rec["index"]
        assert int(out.hash[i]) == rec["hash"], rec["index"]
        assert out.hist_dict(i) == {int(k): v for k, v in rec["hist"].items()}, rec["index"]
    assert out.total_count == gd["total_count"]


@pytest.mark.parametrize("name,golden,tune_stride", [("cl4m-core300", "cl4m_s4_core300.json", 8),
                                                      ("rmat24-core500", "rmat24_s5_core500.json", 16)])
def test_learned_and_tuned_schedule_full_size(name, golden, tune_stride):
    """The path the benchmark times: per-root costs learned by the GPU cost
    model from a sampled plain run, the learned run order, splits and routed
    roots (learned_group_schedule), the packed search with the settings the
    autotuner picks; the results equal the oracle golden of the whole
    workload."""
    gd = load_golden(golden)
    w = workloads.load(name)
    assert w.n_roots == gd["roots"]
    banner(f"{name}: fit cost model")
    pred, model, rep = learned.fit_and_predict(w.graph, w.n_roots)
    model.free()
    assert np.all(np.isfinite(pred))
    v0 = (w.graph.roots(R.ROOT_EDGE) >> np.uint64(32)).astype(np.int64)
    sched = learned.learned_group_schedule(pred, np.arange(w.n_roots, dtype=np.uint64), v0, 4096)
    assert sched.learned and sched.order.size == w.n_roots
    plan = R.Plan(w.graph, root_order=sched.order, eager=sched.eager, runs=sched.runs, pack=True)
    try:
        banner(f"{name}: autotune (stride {tune_stride})")
        tune = plan.autotune(stride=tune_stride)
        assert tune["pack_slots"] in (1, 2, 4, 8) and tune["pack_waves_per_cu"] in (12, 16, 20, 24)
        banner(f"{name}: tuned run {tune}")
        out = plan.run()
    finally:
        plan.free()
    assert out.count == gd["count"]
    assert out.hash == gd["hash"]
    assert out.hist_dict() == {int(k): v for k, v in gd["hist"].items()}
    assert out.total_nodes == gd["tree_nodes"]


def test_batch10k_heavy_sample_packed():
    """The heavy-graph sample of the batch family through the packed search
    in one launch: per-graph count, hash and histogram equal the oracle's."""
    gd = load_golden("batch10k_s3_heavy_sample.json")
    graphs = []
    for rec in gd["graphs"]:
        n, p, s = R.batch_graph_params(3, rec["index"], 32, 1024, 0.05, 0.5)
        g = R.Graph.erdos_renyi(n, p, s)
        g.SortByOrder(R.OrderType.DEG)
        graphs.append(g)
    banner(f"batch10k heavy sample, packed: {len(graphs)} graphs")
    b = R.Batch(graphs, pack=True)
    try:
        out = b.run()
    finally:
        b.free()
    for i, rec in enumerate(gd["graphs"]):
        assert int(out.count[i]) == rec["count"], rec["index"]
        assert int(out.hash[i]) == rec["hash"], rec["index"]
        assert out.hist_dict(i) == {int(k): v for k, v in rec["hist"].items()}, rec["index"]
