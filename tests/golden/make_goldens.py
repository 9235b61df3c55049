"""Writes the committed golden fixtures of tests/golden/ (run from the
repository root, in a container with networkx installed):

  zachary.json         networkx find_cliques of RDMCE/data/zachary.txt
  er2048_p0.008_s1.txt the BASELINE config-1 graph, G(2048, 0.008) seed 1,
                       as COO text (ids as generated, one "u v" line per edge)
  er2048_p0.008_s1.json networkx goldens of that graph
  random_small.json    80 small random graphs (edges) with their goldens

Each golden holds the clique count, the size histogram, the order
independent clique hash (sum over cliques of splitmix64(sum of
splitmix64(label)), mod 2^64) and the SHA-256 of the sorted list of sorted
cliques (original labels).  networkx is independent of both the CPU oracle
and the HIP kernels.
"""
import hashlib
import json
import os
import sys

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import clique_hash  # noqa: E402  (pure-python hash helper)


def golden(edges):
    G = nx.Graph()
    G.add_edges_from((int(a), int(b)) for a, b in edges if a != b)
    cl = [tuple(sorted(c)) for c in nx.find_cliques(G)]
    cl.sort()
    hist = {}
    for c in cl:
        hist[len(c)] = hist.get(len(c), 0) + 1
    sha = hashlib.sha256(json.dumps(cl).encode()).hexdigest()
    return {"count": len(cl), "hist": {str(k): v for k, v in sorted(hist.items())}, "hash": clique_hash(cl),
            "sha256_sorted": sha}


def read_coo(path):
    edges = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 2 and t[0].isdigit() and t[1].isdigit():
                edges.append((int(t[0]), int(t[1])))
    return edges


def main():
    z = read_coo(os.path.join(HERE, "zachary.txt"))
    with open(os.path.join(HERE, "zachary.json"), "w") as f:
        json.dump(golden(z), f, indent=1)

    import rdmce_amd as R
    g = R.Graph.erdos_renyi(2048, 0.008, 1)
    coo = os.path.join(HERE, "er2048_p0.008_s1.txt")
    g.WriteCooText(coo)
    with open(os.path.join(HERE, "er2048_p0.008_s1.json"), "w") as f:
        json.dump(dict(golden(read_coo(coo)), n=2048, p=0.008, seed=1), f, indent=1)

    rng = np.random.default_rng(20261202)
    cases = []
    for i in range(80):
        n = int(rng.integers(5, 61))
        p = float(rng.choice([0.05, 0.1, 0.2, 0.3, 0.5, 0.7, 0.85, 0.95]))
        m = rng.random((n, n)) < p
        iu = np.triu_indices(n, 1)
        sel = m[iu]
        # sparse, shuffled labels exercise the relabelling
        lab = rng.permutation(10 * n)[:n] * 3 + 1
        edges = [(int(lab[a]), int(lab[b])) for a, b in zip(iu[0][sel], iu[1][sel])]
        if not edges:
            continue
        cases.append({"n": n, "p": p, "edges": edges, **golden(edges)})
    with open(os.path.join(HERE, "random_small.json"), "w") as f:
        json.dump(cases, f)
    print("wrote goldens:", len(cases), "random graphs")


if __name__ == "__main__":
    main()
