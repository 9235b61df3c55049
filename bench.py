#!/usr/bin/env python3
"""Maximal clique enumeration on one MI355X: maximal cliques/s.

A "step" is one enumeration pass of the workload by the HIP kernels, with
the graph, the root list and all scratch already resident in HBM (a
device-resident plan); the timed region covers exactly `--steps` passes,
synchronised on both sides.  Prints one JSON line.

Workloads (synthetic, generated on the host by rdmce_amd.workloads):
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
  cl4m-core300    Chung-Lu + planted communities, n = 4e6 (rdmce_amd.workloads).
  rmat24-core500  R-MAT scale 24 (seed 5), roots with core(v0) <= 500
                  (rdmce_amd.workloads).
  batch10k-e1e6   6,315 small G(n, p) graphs in one launch (rdmce_amd.workloads).
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


def read_profile_value(workload: str, kind: str, key: str):
    """A figure from a committed rocprofv3 summary (profiles/), or None:
    kind "pmc" holds traffic_bytes_per_launch (FETCH_SIZE + WRITE_SIZE,
    calibrated), kind "sq" valu_issue_frac and lds_bw_frac."""
    path = os.path.join(ROOT, "profiles", f"mce_kernel_{workload.replace('-', '_')}_{kind}.md")
    try:
        with open(path) as f:
            m = re.search(key + r":\s*([0-9.eE+]+)", f.read())
            return float(m.group(1)) if m else None
    except OSError:
        return None


def read_traffic(workload: str):
    return read_profile_value(workload, "pmc", "traffic_bytes_per_launch")


def golden_check(workload: str, out):
    """Compares a result with the committed oracle golden of `workload`
    (tests/golden/*.json whose "workload" field names it): True on a
    bit-exact match of count, hash and size histogram, None when no golden
    exists; a mismatch stops the run."""
    gdir = os.path.join(ROOT, "tests", "golden")
    for fn in sorted(os.listdir(gdir)):
        if not fn.endswith(".json"):
            continue
        with open(os.path.join(gdir, fn)) as f:
            gd = json.load(f)
        if not isinstance(gd, dict) or gd.get("workload") != workload:
            continue
        hist = {int(k): int(v) for k, v in gd["hist"].items()}
        if (out.count, out.hash, out.hist_dict()) != (gd["count"], gd["hash"], hist):
            raise SystemExit(f"{workload}: results differ from the committed oracle golden {fn}")
        return True
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


def time_plan(plan, steps: int, warmup: int, label: str):
    """Runs `warmup` untimed and `steps` timed passes of a resident plan;
    returns (outputs, wall seconds of the timed passes)."""
    import numpy as np

    from rdmce_amd._lib import check, lib

    for i in range(warmup):
        o = plan.run()
        print(f"[bench] {label} warmup {i}: {o.count} cliques, kernel {o.kernel_ms:.1f} ms", file=sys.stderr,
              flush=True)
    check(lib().mce_device_sync(0))
    t0 = time.perf_counter()
    outs = []
    for i in range(steps):
        outs.append(plan.run())
        print(f"[bench] {label} step {i}: kernel {outs[-1].kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    check(lib().mce_device_sync(0))
    wall = time.perf_counter() - t0
    first = outs[0]
    for o in outs[1:]:
        if (o.count, o.hash) != (first.count, first.hash) or not np.array_equal(o.hist, first.hist):
            raise SystemExit(f"{label} steps disagree: results are not reproducible")
    return outs, wall


def run_batch_workload(args):
    """Batch workloads: all graphs in one persistent launch (a replayed
    hipGraph) per step; the plain path solves one graph per call, as the
    reference handles one graph per invocation (every 10th graph, timed)."""
    import numpy as np
    import rdmce_amd as R
    from rdmce_amd import workloads

    t = time.perf_counter()
    bw = workloads.load_batch(args.workload)
    prep_s = time.perf_counter() - t
    print(f"[bench] {bw.desc} prepared in {prep_s:.1f}s", file=sys.stderr, flush=True)
    golden = os.path.join(ROOT, "tests", "golden", "batch10k_s3_" + args.workload.split("-")[1] + ".json")
    b = R.Batch(bw.graphs, use_graph=True)
    for i in range(args.warmup):
        o = b.run()
        print(f"[bench] batch warmup {i}: {o.total_count} cliques, kernel {o.kernel_ms:.1f} ms", file=sys.stderr,
              flush=True)
    t0 = time.perf_counter()
    outs = []
    for i in range(args.steps):
        outs.append(b.run())
        print(f"[bench] batch step {i}: kernel {outs[-1].kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    b.free()
    first = outs[0]
    for o in outs[1:]:
        if not (np.array_equal(o.count, first.count) and np.array_equal(o.hash, first.hash)):
            raise SystemExit("batch steps disagree")
    digest = workloads.batch_digest(first.count, first.hash)
    check = None
    if os.path.exists(golden):
        with open(golden) as f:
            gd = json.load(f)
        check = bool(gd["digest"] == digest and gd["total_count"] == first.total_count)
        if not check:
            raise SystemExit("batch results differ from the committed oracle golden")
    # plain: one call per graph on every 10th graph
    sample = bw.graphs[::10]
    t = time.perf_counter()
    sc = 0
    for g in sample:
        sc += R.solve(g).count
    plain_wall = time.perf_counter() - t
    value = first.total_count * args.steps / wall
    line = {
        "metric": METRIC, "value": value, "unit": "maximal cliques/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": dict(bw.desc, cliques=first.total_count, tree_nodes=first.total_nodes, digest=digest,
                       golden_match=check, kernel_ms=sum(o.kernel_ms for o in outs) / len(outs),
                       graph_replay=first.graph_replay, passes=first.passes, prep_s=round(prep_s, 2), mode="batch"),
        "plain": {"value": sc / plain_wall, "sample": f"every 10th graph ({len(sample)} graphs, {sc} cliques), "
                  f"one mce_solve call each, {plain_wall:.2f} s"},
        "speedup_vs_plain": value / (sc / plain_wall),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat20-core499")
    ap.add_argument("--mode", choices=["tuned", "plain"], default="tuned",
                    help="tuned: time the plain path, then the autotuned settings with the learned schedule "
                         "or, when one untimed run of each shows it faster, the plain order (value); "
                         "plain: the plain path only")
    ap.add_argument("--share-bound", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--tune-stride", type=int, default=1, help="autotune on every n-th root (1: all)")
    ap.add_argument("--heavy-fraction", type=float, default=0.01,
                    help="learned schedule: share of roots dispatched first by predicted cost (1: all)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=2,
                    help="time the CPU oracle on every n-th root of the workload (2: ~15 s on 16 threads)")
    ap.add_argument("--copy-probe", action="store_true", help="also launch the traffic calibration copy kernel")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload.startswith("batch"):
        return run_batch_workload(args)

    import numpy as np
    import rdmce_amd as R
    from rdmce_amd import learned
    from rdmce_amd._lib import check, lib

    t_prep = time.perf_counter()
    g, ids, desc = build_workload(args.workload)
    prep_s = time.perf_counter() - t_prep
    n_roots = int(ids.size)
    print(f"[bench] workload {desc} prepared in {prep_s:.1f}s", file=sys.stderr, flush=True)

    # plain path: reverse canonical order, default settings (the reference's dispatch)
    plan = R.Plan(g, root_order=ids, share_bound=args.share_bound, waves_per_cu=args.waves_per_cu)
    plain_outs, plain_wall = time_plan(plan, args.steps, args.warmup, "plain")
    plan.free()
    plain = plain_outs[0]
    plain_info = {"value": plain.count * args.steps / plain_wall, "ms_per_step": plain_wall * 1e3 / args.steps,
                  "kernel_ms": sum(o.kernel_ms for o in plain_outs) / len(plain_outs),
                  "schedule": "reverse canonical order (high core first)", "waves_per_cu": 16,
                  "share_bound": args.share_bound or 24}

    outs, wall, sched_info = plain_outs, plain_wall, {"schedule": "plain"}
    if args.mode == "tuned":
        t = time.perf_counter()
        pred, model, rep = learned.fit_and_predict(g, n_roots)
        sched = learned.learned_schedule(pred, np.arange(n_roots, dtype=np.uint64),
                                         heavy_fraction=args.heavy_fraction)
        plan = R.Plan(g, root_order=sched.order, eager=sched.eager)
        tune = plan.autotune(stride=args.tune_stride)
        # The plain order with the tuned settings is the fallback: one untimed
        # run of each order, and the faster one is timed.
        alt = R.Plan(g, root_order=ids, waves_per_cu=tune["waves_per_cu"], share_bound=tune["share_bound"],
                     lds_tier=tune["lds_tier"])
        trial = {"learned": plan.run().kernel_ms, "plain": alt.run().kernel_ms}
        use_learned = bool(sched.learned) and trial["learned"] <= trial["plain"]
        if use_learned:
            alt.free()
        else:
            plan.free()
            plan = alt
        prep2 = time.perf_counter() - t
        print(f"[bench] learned+tuned prep {prep2:.1f}s: fit {rep}, tune {tune}, order trial ms {trial}",
              file=sys.stderr, flush=True)
        outs, wall = time_plan(plan, args.steps, args.warmup, "tuned")
        plan.free()
        o = outs[0]
        if (o.count, o.hash) != (plain.count, plain.hash) or not np.array_equal(o.hist, plain.hist):
            raise SystemExit("learned+tuned results differ from the plain path")
        sched_info = {"schedule": "learned (predicted-heaviest first, then plain order)" if use_learned
                      else "plain order with the tuned settings (measured faster than the learned order)",
                      "learned": use_learned, "order_trial_ms": trial,
                      "model_loss": rep.loss, "label_roots": rep.sample_roots,
                      "heavy_fraction": args.heavy_fraction,
                      "waves_per_cu": tune["waves_per_cu"], "share_bound": tune["share_bound"],
                      "lds_tier": tune["lds_tier"], "tune_trials": tune["trials"],
                      "tune_stride": args.tune_stride, "learn_tune_s": round(prep2, 2)}

    first = outs[0]
    golden_match = golden_check(args.workload, first)
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
                       waves=first.waves, golden_match=golden_match, prep_s=round(prep_s, 2), mode=args.mode,
                       **sched_info),
        "plain": plain_info,
        "speedup_vs_plain": value / plain_info["value"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(args.workload),
                     "algo_bytes": algo_bytes,
                     "valu_issue_frac": read_profile_value(args.workload, "sq", "valu_issue_frac"),
                     "lds_bw_frac": read_profile_value(args.workload, "sq", "lds_bw_frac")},
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(g, ids, args.cpu_stride, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
