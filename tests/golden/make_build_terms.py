"""Writes tests/golden/build_terms.json: per workload, the algorithmic build
terms of its edge roots as the oracle builds them (d(v0) + d(v1), the
degrees of P0's members, rows x words of each root's bitmap) and the build
bytes they give, 4 deg + 4 pdeg + 8 rows_w.  One-off (R-MAT 24 takes some
minutes on 8 threads):

    python tests/golden/make_build_terms.py rmat20-core499 cl4m-core300 rmat24-core500
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle_lib import OracleGraph  # noqa: E402
from rdmce_amd import workloads  # noqa: E402

OUT = os.path.join(HERE, "build_terms.json")


def main():
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in sys.argv[1:]:
        w = workloads.load(name)
        g = w.graph
        og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
        t = time.time()
        terms = og.build_terms(0, w.n_roots)
        terms["seconds"] = round(time.time() - t, 1)
        data[name] = terms
        print(name, terms, flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
