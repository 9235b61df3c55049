#!/usr/bin/env python3
"""Maximal clique enumeration on one MI355X: maximal cliques/s.

A "step" is one enumeration pass of the workload by the HIP kernels, with
the graph, the root list and all scratch already resident in HBM (a
device-resident plan); the timed region covers exactly `--steps` passes,
synchronised on both sides.  Prints one JSON line.

Workloads (synthetic, generated on the host):
  rmat20-core499  R-MAT scale 20, edge factor 16, (a,b,c,d) =
                  (0.57,0.19,0.19,0.05), seed 2, symmetrised, deduplicated,
                  no self loops, CSR in degeneracy order.  Enumerated: every
                  maximal clique whose two lowest vertices in degeneracy
                  order (the edge root v0 < v1) have core(v0) <= 499, i.e.
                  13,226,611 of the 15,700,316 edge roots and 1,035,193,387
                  cliques.  The roots above core 499 hold a dense core whose
                  clique count grows by >1e9 per core level (full enumeration
                  is out of reach for any implementation).
  er2048          Erdos-Renyi G(2048, 0.008), seed 1 (all roots).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MCE wall time & maximal cliques/s, 1 MI355X; batched+learned vs plain"


def build_workload(name: str):
    from rdmce_amd import workloads

    w = workloads.load(name)
    return w.graph, w.plain_order(), w.desc


def read_traffic(workload: str):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE
    (committed summary), or None."""
    path = os.path.join(ROOT, "profiles", f"mce_kernel_{workload.replace('-', '_')}_pmc.md")
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
            return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(g, ids, stride: int, threads: int):
    """CPU oracle (multithreaded C port of the reference algorithm) on every
    `stride`-th root of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleGraph
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    sample = ids[::stride]
    t0 = time.perf_counter()
    r = og.mce_roots(sample, root_mode=2, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": r.count / dt, "unit": "maximal cliques/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}th root of the workload ({sample.size} roots, {r.count} cliques) "
                      f"in {dt:.2f} s on {threads} host threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat20-core499")
    ap.add_argument("--share-bound", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--copy-probe", action="store_true", help="also launch the traffic calibration copy kernel")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import rdmce_amd as R
    from rdmce_amd._lib import check, lib

    t_prep = time.perf_counter()
    g, ids, desc = build_workload(args.workload)
    prep_s = time.perf_counter() - t_prep
    print(f"[bench] workload {desc} prepared in {prep_s:.1f}s", file=sys.stderr, flush=True)

    plan = R.Plan(g, root_order=ids, share_bound=args.share_bound, waves_per_cu=args.waves_per_cu)
    for i in range(args.warmup):
        o = plan.run()
        print(f"[bench] warmup {i}: {o.count} cliques, kernel {o.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    check(lib().mce_device_sync(0))
    t0 = time.perf_counter()
    outs = []
    for i in range(args.steps):
        outs.append(plan.run())
        print(f"[bench] step {i}: kernel {outs[-1].kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    check(lib().mce_device_sync(0))
    t1 = time.perf_counter()
    plan.free()
    first = outs[0]
    for o in outs[1:]:
        if (o.count, o.hash) != (first.count, first.hash) or not np.array_equal(o.hist, first.hist):
            raise SystemExit("steps disagree: results are not reproducible")

    wall = t1 - t0
    value = first.count * args.steps / wall
    kernel_ms = sum(o.kernel_ms for o in outs) / len(outs)
    algo_bytes = first.search_bytes + first.build_bytes
    achieved = algo_bytes / (kernel_ms / 1e3) / 1e9
    if args.copy_probe:
        import ctypes
        ms = ctypes.c_double()
        check(lib().mce_copy_probe(0, 1 << 30, 4, ctypes.byref(ms)))
        print(f"[bench] copy probe 1 GiB: {ms.value:.3f} ms", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "maximal cliques/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": dict(desc, cliques=first.count, clique_hash=first.hash, tree_nodes=first.total_nodes,
                       kernel_ms=kernel_ms, shared_tasks=first.shared_tasks, passes=first.passes,
                       waves=first.waves, schedule="plain (reverse root order)", prep_s=round(prep_s, 2)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(args.workload),
                     "algo_bytes": algo_bytes},
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(g, ids, args.cpu_stride, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
