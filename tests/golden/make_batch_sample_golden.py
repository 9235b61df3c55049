"""Writes the CPU-oracle golden of a stratified sample of the heavy graphs of
the 10,000-graph batch family (G(n, p), n uniform in [32, 1024], p uniform
in [0.05, 0.5], seed 3): the graphs whose closed-form expected number of
maximal cliques exceeds 1e6, which the batch10k-e1e6 workload leaves out.

The sample is deterministic: the excluded graphs are grouped by decade of
their expected count (1e6..1e7, ..., >= 1e10), sorted by (n, p) within a
decade, and evenly spaced graphs are taken from each group (PER_DECADE).
For every sampled graph the golden holds its family index, n, p, count,
clique hash, size histogram and search-tree nodes.  Run from the
repository root (about 20 minutes on 8 host threads):

  python tests/golden/make_batch_sample_golden.py tests/golden/batch10k_s3_heavy_sample.json
"""
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import rdmce_amd as R  # noqa: E402
from rdmce_amd.workloads import expected_er_cliques  # noqa: E402
from oracle_lib import OracleGraph  # noqa: E402

PER_DECADE = {6: 8, 7: 8, 8: 8, 9: 4, 10: 1}


def sample_indices():
    by_decade = {}
    for i in range(10000):
        n, p, s = R.batch_graph_params(3, i, 32, 1024, 0.05, 0.5)
        e = expected_er_cliques(n, p)
        if e <= 1e6:
            continue
        by_decade.setdefault(min(int(math.log10(e)), 10), []).append((n, p, i))
    picked = []
    for d, k in sorted(PER_DECADE.items()):
        grp = sorted(by_decade.get(d, []))
        if not grp:
            continue
        step = len(grp) / k
        picked += [grp[int(j * step + step / 2)][2] for j in range(min(k, len(grp)))]
    return sorted(picked)


def main():
    out = sys.argv[1]
    recs = []
    t0 = time.time()
    for i in sample_indices():
        n, p, s = R.batch_graph_params(3, i, 32, 1024, 0.05, 0.5)
        g = R.Graph.erdos_renyi(n, p, s)
        g.SortByOrder(R.OrderType.DEG)
        og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
        t = time.time()
        r = og.mce()
        recs.append({"index": i, "n": n, "p": p, "expected": expected_er_cliques(n, p), "count": r.count,
                     "hash": r.hash, "hist": {str(k): v for k, v in sorted(r.hist.items())},
                     "tree_nodes": r.total_nodes, "oracle_seconds": round(time.time() - t, 2)})
        print(json.dumps({k: recs[-1][k] for k in ("index", "n", "p", "count", "oracle_seconds")}), flush=True)
    rec = {"family": "G(n, p), n ~ U[32,1024], p ~ U[0.05,0.5], seed 3, degeneracy order",
           "rule": "stratified sample of the graphs with expected maximal cliques > 1e6",
           "per_decade": {str(k): v for k, v in PER_DECADE.items()}, "graphs": recs,
           "total_count": sum(r["count"] for r in recs), "oracle_seconds": round(time.time() - t0, 1),
           "oracle_threads": os.cpu_count()}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({"graphs": len(recs), "total_count": rec["total_count"], "seconds": rec["oracle_seconds"]}))


if __name__ == "__main__":
    main()
