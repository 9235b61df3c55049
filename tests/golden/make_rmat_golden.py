"""Writes the CPU-oracle golden (count, histogram, clique hash, tree nodes)
of an R-MAT workload: all edge roots (v0 < v1 in degeneracy order) with
core(v0) <= CORE_MAX.  Run from the repository root; takes minutes on 8
host threads for rmat20 core 499.

  python tests/golden/make_rmat_golden.py 20 2 499 tests/golden/rmat20_s2_core499.json
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import rdmce_amd as R  # noqa: E402
from oracle_lib import OracleGraph  # noqa: E402


def main():
    scale, seed, core_max, out = int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3]), sys.argv[4]
    g = R.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=seed)
    g.SortByOrder(R.OrderType.DEG)
    core = g.GetCoreNumbers()
    roots = g.roots(R.ROOT_EDGE)
    v0 = (roots >> np.uint64(32)).astype(np.int64)
    end = int(np.searchsorted(core[v0], core_max, side="right"))
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    t = time.time()
    r = og.mce(root_begin=0, root_end=end)
    rec = {"graph": f"rmat scale {scale} ef 16 (0.57,0.19,0.19,0.05) seed {seed}, degeneracy order",
           "n": g.GetNumVertices(), "m": g.GetNumEdges(), "degeneracy": g.GetDegeneracy(),
           "core_max": core_max, "roots": end, "count": r.count, "hash": r.hash,
           "hist": {str(k): v for k, v in sorted(r.hist.items())}, "tree_nodes": r.total_nodes,
           "oracle_seconds": round(time.time() - t, 1), "oracle_threads": os.cpu_count()}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({k: rec[k] for k in ("roots", "count", "hash", "tree_nodes", "oracle_seconds")}))


if __name__ == "__main__":
    main()
