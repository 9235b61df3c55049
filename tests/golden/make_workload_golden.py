"""Writes the CPU-oracle golden (count, histogram, clique hash, tree nodes)
of a named core-bounded workload (rdmce_amd.workloads): all edge roots
(v0 < v1 in degeneracy order) with core(v0) <= c.  Run from the repository
root; minutes on 8 host threads.

  python tests/golden/make_workload_golden.py rmat20-core499 tests/golden/rmat20_s2_core499.json
  python tests/golden/make_workload_golden.py cl4m-core300 tests/golden/cl4m_s4_core300.json
  python tests/golden/make_workload_golden.py rmat24-core500 tests/golden/rmat24_s5_core500.json
  python tests/golden/make_workload_golden.py batch10k-e1e6 tests/golden/batch10k_s3_e1e6.json

A batch golden holds the totals and a digest of the per-graph (count, hash)
pairs (rdmce_amd.workloads.batch_digest).
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from rdmce_amd import workloads  # noqa: E402
from oracle_lib import OracleGraph  # noqa: E402


def batch_golden(name, out):
    bw = workloads.load_batch(name)
    t = time.time()
    counts, hashes, hist, nodes = [], [], {}, 0
    for g in bw.graphs:
        og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
        r = og.mce()
        counts.append(r.count)
        hashes.append(r.hash)
        nodes += r.total_nodes
        for k, v in r.hist.items():
            hist[k] = hist.get(k, 0) + v
    rec = {"workload": name, "graph": bw.desc["graph"], "rule": bw.desc["rule"], "graphs": len(bw.graphs),
           "total_count": sum(counts), "digest": workloads.batch_digest(counts, hashes),
           "hist": {str(k): v for k, v in sorted(hist.items())}, "tree_nodes": nodes,
           "oracle_seconds": round(time.time() - t, 1), "oracle_threads": os.cpu_count()}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({k: rec[k] for k in ("graphs", "total_count", "digest", "tree_nodes", "oracle_seconds")}))


def main():
    name, out = sys.argv[1], sys.argv[2]
    if name.startswith("batch"):
        return batch_golden(name, out)
    w = workloads.load(name)
    g = w.graph
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    t = time.time()
    r = og.mce(root_begin=0, root_end=w.n_roots)
    rec = {"workload": name, "graph": w.desc["graph"],
           "n": g.GetNumVertices(), "m": g.GetNumEdges(), "degeneracy": g.GetDegeneracy(),
           "core_max": int(name.rsplit("core", 1)[1]), "roots": w.n_roots, "count": r.count, "hash": r.hash,
           "hist": {str(k): v for k, v in sorted(r.hist.items())}, "tree_nodes": r.total_nodes,
           "oracle_seconds": round(time.time() - t, 1), "oracle_threads": os.cpu_count()}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({k: rec[k] for k in ("roots", "count", "hash", "tree_nodes", "oracle_seconds")}))


if __name__ == "__main__":
    main()
