#!/usr/bin/env python3
"""Maximal clique enumeration on one MI355X: maximal cliques/s, batched +
learned against plain.

A "step" is one enumeration pass of the workload by the HIP kernels over a
device-resident plan (graph, root list, schedule and scratch already in
HBM); each timed leg covers exactly its steps, synchronised on both sides.
Prints one JSON line.

Legs (identical results, checked bit for bit against each other and against
the committed CPU-oracle golden of the workload):
  plain      the one-root-per-wavefront kernel in the reference's dispatch
             order (high core first, BkPivotBitBalance.cuh:429-431) with its
             default settings (16 waves per CU, sharing bound 24);
  batched + learned (value)
             the packed search - many roots' subproblems resident per
             wavefront, one search-tree node per lane, the edge roots of a
             vertex derived from one shared bitmap - with the dispatch order
             and run splits of the learned cost model (GPU-trained FP32 MLP)
             and the settings of the autotuner;
  batched, plain order
             the same packed search and settings in the plain order: what the
             learned schedule adds.

Workloads (synthetic, generated on the host by rdmce_amd.workloads):
  rmat24-core500  R-MAT scale 24, edge factor 16, (a,b,c,d) =
                  (0.57,0.19,0.19,0.05), seed 5, symmetrised, deduplicated,
                  no self loops, CSR in degeneracy order (8.9M vertices, 260M
                  edges).  Enumerated: every maximal clique whose two lowest
                  vertices (v0 < v1 in degeneracy order) have core(v0) <= 500
                  - 210.7M of the 260.4M edge roots, 1.02e10 cliques.
  cl4m-core300    Chung-Lu + planted communities, n = 4e6, core(v0) <= 300.
  rmat20-core499  R-MAT scale 20, seed 2, core(v0) <= 499.
  er2048          Erdos-Renyi G(2048, 0.008), seed 1 (all roots).
  batch10k-e1e6   6,315 small G(n, p) graphs in one launch.
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

METRIC = "MCE wall time & maximal cliques/s, 1 MI355X; batched+learned vs plain"
# MI355X peaks for integer set work (no MFMA): HBM3E, LDS (128 B/clk/CU x
# 256 CUs x 2.4 GHz) and 32-bit VALU (64 lanes/clk/CU x 256 CUs x 2.4 GHz).
HBM_PEAK_GBS = 8000.0
LDS_PEAK_GBS = 78643.2
VALU_PEAK_GOPS = 39321.6
OPS_PER_WORD = 6  # per 64-bit bitmap word: ANDN, POPCOUNT (two 32-bit halves each), compare/accumulate
BOX_THREADS = 16  # the host threads of one GPU's share of the box


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_profile(name: str, key: str):
    """A figure from a committed rocprofv3 summary (profiles/<name>.md), or
    None."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    try:
        with open(path) as f:
            m = re.search(re.escape(key) + r":\s*([0-9.eE+]+)", f.read())
            return float(m.group(1)) if m else None
    except OSError:
        return None


def golden_for(workload: str):
    gdir = os.path.join(ROOT, "tests", "golden")
    for fn in sorted(os.listdir(gdir)):
        if fn.endswith(".json"):
            with open(os.path.join(gdir, fn)) as f:
                gd = json.load(f)
            if isinstance(gd, dict) and gd.get("workload") == workload:
                return fn, gd
    return None, None


def check_golden(workload: str, out):
    """True on a bit-exact match of count, hash and size histogram with the
    committed oracle golden, None without one; a mismatch stops the run."""
    fn, gd = golden_for(workload)
    if gd is None:
        return None
    hist = {int(k): int(v) for k, v in gd["hist"].items()}
    if (out.count, out.hash, out.hist_dict()) != (gd["count"], gd["hash"], hist):
        raise SystemExit(f"{workload}: results differ from the committed oracle golden {fn}")
    return True


def time_plan(plan, steps: int, warmup: int, label: str):
    """`warmup` untimed and `steps` timed passes of a resident plan; returns
    (outputs, wall seconds of the timed passes)."""
    import numpy as np

    from rdmce_amd._lib import check, lib

    for i in range(warmup):
        o = plan.run()
        log(f"{label} warmup {i}: {o.count} cliques, kernel {o.kernel_ms:.1f} ms")
    check(lib().mce_device_sync(0))
    t0 = time.perf_counter()
    outs = []
    for i in range(steps):
        outs.append(plan.run())
        log(f"{label} step {i}: kernel {outs[-1].kernel_ms:.1f} ms")
    check(lib().mce_device_sync(0))
    wall = time.perf_counter() - t0
    first = outs[0]
    for o in outs[1:]:
        if (o.count, o.hash) != (first.count, first.hash) or not np.array_equal(o.hist, first.hist):
            raise SystemExit(f"{label} steps disagree: results are not reproducible")
    return outs, wall


def same(a, b) -> bool:
    import numpy as np
    return (a.count, a.hash, a.total_nodes) == (b.count, b.hash, b.total_nodes) and np.array_equal(a.hist, b.hist)


def roofline(out, kernel_ms: float, workload: str, plain=None):
    """Three ceilings of this integer set work (no MFMA): the subproblem
    builds' bytes against HBM, the search's bitmap words against LDS, and
    the search's word operations against VALU.  T_lb is the largest of the
    three times; frac = T_lb / measured kernel time, and `bound` names the
    binding ceiling, whose achieved rate and peak are reported.

    Build bytes per root, one denominator for every schedule of the same
    roots: the bytes of building each root on its own, as the
    one-root-per-wave kernel (`plain`, the same roots) counts them - 4 B per
    adjacency entry its lists read, 4 B per entry of each candidate's
    out-list run matched, 4 B per list entry, 8 B per bitmap word.  The
    search bytes and operations come from the canonical tree, the same for
    every kernel and order.  Beside them: the bytes the measured kernels
    count for their own builds (the packed search builds one bitmap per
    first vertex, so it moves fewer), and the per-root formula with whole
    adjacency lists, summed by the oracle."""
    t = kernel_ms / 1e3
    b_own = out.build_bytes
    b_build = plain.build_bytes if plain is not None else b_own
    b_search = out.search_bytes
    ops = OPS_PER_WORD * (b_search // 8)
    ceil = {"hbm": (b_build / HBM_PEAK_GBS / 1e9, b_build / t / 1e9, HBM_PEAK_GBS, "GB/s"),
            "lds": (b_search / LDS_PEAK_GBS / 1e9, b_search / t / 1e9, LDS_PEAK_GBS, "GB/s"),
            "valu": (ops / VALU_PEAK_GOPS / 1e9, ops / t / 1e9, VALU_PEAK_GOPS, "Gop/s")}
    bound = max(ceil, key=lambda k: ceil[k][0])
    t_lb, achieved, peak, unit = ceil[bound]
    own_t_lb = max(b_own / HBM_PEAK_GBS / 1e9, ceil["lds"][0], ceil["valu"][0])
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": t_lb / t,
            "traffic": read_profile(f"pack_{workload.replace('-', '_')}_pmc", "traffic_bytes_per_step"),
            "t_lb_ms": t_lb * 1e3, "build_bytes": b_build,
            "build_bytes_denominator": "per-root builds (one-root-per-wave kernel, same roots)"
            if plain is not None else "the measured kernels' own builds",
            "search_bytes": b_search, "search_ops": ops,
            "ceilings_ms": {k: v[0] * 1e3 for k, v in ceil.items()},
            "frac_of_peak": {k: v[1] / v[2] for k, v in ceil.items()},
            "own_build": {"build_bytes": b_own, "hbm_ms": b_own / HBM_PEAK_GBS / 1e6, "frac": own_t_lb / t},
            "per_root_build_formula": per_root_build(workload, t)}


def per_root_build(workload: str, t: float):
    """The per-root build formula 4 (d(v0) + d(v1)) + 4 sum_{u in P0} d(u) +
    8 rows W, summed over the workload's roots by the oracle
    (tests/golden/build_terms.json): the bytes a build that scans each
    candidate's whole adjacency list for every root would move, and the HBM
    time they would take at peak beside the measured step."""
    fn = os.path.join(ROOT, "tests", "golden", "build_terms.json")
    try:
        with open(fn) as f:
            terms = json.load(f).get(workload)
    except OSError:
        terms = None
    if not terms:
        return None
    hbm_s = terms["build_bytes"] / (HBM_PEAK_GBS * 1e9)
    return {"build_bytes": terms["build_bytes"], "hbm_ms": hbm_s * 1e3, "over_measured": hbm_s / t}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, n_roots: int, seconds: float, prep_s: float):
    """The CPU oracle (multithreaded C restatement of the reference
    algorithm) on a bounded sample of the workload's roots - every k-th
    root, k chosen from a short calibration run so that the sample takes
    about `seconds` - on this host's threads (the GPU's share of the box),
    plus its one-thread rate on the ER G(2048, 0.008) graph."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleGraph
    import rdmce_amd as R

    threads = min(BOX_THREADS, len(os.sched_getaffinity(0)))
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    ids = np.arange(n_roots, dtype=np.uint64)[::-1].copy()
    probe = ids[::max(1, n_roots // 20000)]
    t0 = time.perf_counter()
    og.mce_roots(probe, root_mode=2, threads=threads)
    per_root = (time.perf_counter() - t0) / max(probe.size, 1)
    stride = max(1, int(n_roots * per_root / seconds))
    sample = ids[::stride]
    t0 = time.perf_counter()
    r = og.mce_roots(sample, root_mode=2, threads=threads)
    dt = time.perf_counter() - t0
    e = R.Graph.erdos_renyi(2048, 0.008, 1)
    e.SortByOrder(R.OrderType.DEG)
    oe = OracleGraph.from_csr(e.GetRowOffset(), e.GetColIdx(), e.GetLabels(), e.GetDegeneracy())
    t1 = time.perf_counter()
    r1 = oe.mce(threads=1)
    dt1 = time.perf_counter() - t1
    return {"value": r.count / dt, "unit": "maximal cliques/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}th root of the workload ({sample.size} roots, {r.count} cliques, "
                      f"{r.total_nodes} tree nodes) in {dt:.2f} s on {threads} host threads",
            "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": cpu_model(),
            "prep_s": round(prep_s, 2),
            "one_thread_er2048": {"cliques_per_s": r1.count / dt1, "cliques": r1.count, "seconds": dt1}}


def run_batch_workload(args):
    """Batch workloads: the batched leg runs all graphs in one persistent
    launch of the packed search, replayed as a captured hipGraph; the plain
    leg runs one launch per graph (the reference solves one graph per
    invocation, BkPivotBitBalance.cuh:661) over device-resident plans, on
    every 10th graph, timing each launch."""
    import numpy as np
    import rdmce_amd as R
    from rdmce_amd import workloads

    t = time.perf_counter()
    bw = workloads.load_batch(args.workload)
    prep_s = time.perf_counter() - t
    log(f"{bw.desc} prepared in {prep_s:.1f}s")
    b = R.Batch(bw.graphs, use_graph=True, pack=True)
    for i in range(args.warmup):
        o = b.run()
        log(f"batch warmup {i}: {o.total_count} cliques, kernel {o.kernel_ms:.1f} ms")
    t0 = time.perf_counter()
    outs = []
    for i in range(args.steps):
        outs.append(b.run())
        log(f"batch step {i}: kernel {outs[-1].kernel_ms:.1f} ms")
    wall = time.perf_counter() - t0
    b.free()
    first = outs[0]
    for o in outs[1:]:
        if not (np.array_equal(o.count, first.count) and np.array_equal(o.hash, first.hash)):
            raise SystemExit("batch steps disagree")
    digest = workloads.batch_digest(first.count, first.hash)
    golden = os.path.join(ROOT, "tests", "golden", "batch10k_s3_" + args.workload.split("-")[1] + ".json")
    check = None
    if os.path.exists(golden):
        with open(golden) as f:
            gd = json.load(f)
        check = bool(gd["digest"] == digest and gd["total_count"] == first.total_count)
        if not check:
            raise SystemExit("batch results differ from the committed oracle golden")
    sample = bw.graphs[::10]
    sc, kms = 0, 0.0
    for g in sample:
        plan = R.Plan(g)
        plan.run()
        o = plan.run()
        plan.free()
        sc += o.count
        kms += o.kernel_ms
    value = first.total_count * args.steps / wall
    plain_rate = sc / (kms / 1e3)
    line = {
        "metric": METRIC, "value": value, "unit": "maximal cliques/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": dict(bw.desc, cliques=first.total_count, tree_nodes=first.total_nodes, digest=digest,
                       golden_match=check, kernel_ms=sum(o.kernel_ms for o in outs) / len(outs),
                       graph_replay=first.graph_replay, prep_s=round(prep_s, 2), mode="batch, packed search"),
        "plain": {"value": plain_rate, "sample": f"every 10th graph ({len(sample)} graphs, {sc} cliques), one "
                  f"launch per graph on resident plans, {kms:.1f} ms of kernel time"},
        "speedup_vs_plain": value / plain_rate,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat24-core500")
    ap.add_argument("--plain-steps", type=int, default=3, help="timed steps of the plain leg (at most --steps)")
    ap.add_argument("--ablation-steps", type=int, default=3, help="timed steps of the batched plain-order leg")
    ap.add_argument("--tune-stride", type=int, default=8, help="autotune on every n-th run of edge roots")
    ap.add_argument("--chunk-share", type=float, default=0.25)
    ap.add_argument("--route-share", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also launch a 1 GiB streaming copy kernel (traffic counter calibration)")
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
    from rdmce_amd import learned, workloads

    t0 = time.perf_counter()
    w = workloads.load(args.workload)
    prep_s = time.perf_counter() - t0
    g = w.graph
    total_roots = g.GetNumEdges()
    log(f"workload {w.desc} prepared in {prep_s:.1f}s")

    # Plain leg.
    plan = R.Plan(g, root_order=w.plain_order())
    p_steps = max(1, min(args.plain_steps, args.steps))
    plain_outs, plain_wall = time_plan(plan, p_steps, 1, "plain")
    plan.free()
    plain = plain_outs[0]
    golden = check_golden(args.workload, plain)
    plain_info = {"value": plain.count * p_steps / plain_wall, "ms_per_step": plain_wall * 1e3 / p_steps,
                  "kernel_ms": sum(o.kernel_ms for o in plain_outs) / p_steps, "steps": p_steps,
                  "schedule": "reverse canonical order (high core first)", "kernel": "one root per wavefront",
                  "waves_per_cu": 16, "share_bound": 24}

    # Learned schedule: per-root costs from a sampled plain run, the GPU-trained model, its predictions.
    t1 = time.perf_counter()
    pred, model, rep = learned.fit_and_predict(g, w.n_roots)
    model.free()
    ids = np.arange(w.n_roots, dtype=np.uint64)
    v0 = (g.roots(R.ROOT_EDGE) >> np.uint64(32)).astype(np.int64)
    pack_waves = 256 * 16
    sched = learned.learned_group_schedule(pred, ids, v0, pack_waves, chunk_share=args.chunk_share,
                                           route_share=args.route_share)
    learn_s = time.perf_counter() - t1
    log(f"learned schedule in {learn_s:.1f}s: {rep}, {sched.runs.size - 1} runs, {int(sched.eager.sum())} roots "
        f"routed to the splitting kernel")

    # Batched + learned leg (value): autotuned packed search.
    plan = R.Plan(g, root_order=sched.order, eager=sched.eager, runs=sched.runs, pack=True)
    t2 = time.perf_counter()
    tune = plan.autotune(stride=args.tune_stride)
    tune_s = time.perf_counter() - t2
    log(f"autotune {tune_s:.1f}s: {tune}")
    outs, wall = time_plan(plan, args.steps, args.warmup, "batched+learned")
    plan.free()
    first = outs[0]
    if not same(first, plain):
        raise SystemExit("batched+learned results differ from the plain path")
    check_golden(args.workload, first)
    kernel_ms = sum(o.kernel_ms for o in outs) / len(outs)

    # Ablation: the same packed search and settings in the plain order.
    plan = R.Plan(g, root_order=w.plain_order(), pack=True, pack_slots=tune["pack_slots"],
                  pack_waves_per_cu=tune["pack_waves_per_cu"], share_bound=tune["share_bound"])
    a_steps = max(1, min(args.ablation_steps, args.steps))
    abl_outs, abl_wall = time_plan(plan, a_steps, 1, "batched plain order")
    plan.free()
    if not same(abl_outs[0], plain):
        raise SystemExit("batched plain-order results differ from the plain path")

    # Root granularity: vertex roots against the edge roots of the same
    # vertices (every 64th vertex of the covered ones), tuned packed settings.
    nv = int(v0[w.n_roots - 1]) + 1 if w.n_roots else 0
    gran = R.choose_root_mode(g, n_vertex_roots=nv, stride=64, pack=True, pack_slots=tune["pack_slots"],
                              pack_waves_per_cu=tune["pack_waves_per_cu"], share_bound=tune["share_bound"])
    gran["root_mode"] = "vertex" if gran["root_mode"] == R.ROOT_VERTEX else "edge"
    log(f"root granularity: {gran}")

    if args.copy_probe:
        import ctypes
        from rdmce_amd._lib import check, lib
        ms = ctypes.c_double()
        check(lib().mce_copy_probe(0, 1 << 30, 4, ctypes.byref(ms)))
        log(f"copy probe 1 GiB: {ms.value:.3f} ms")

    value = first.count * args.steps / wall
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
        "config": dict(w.desc, cliques=first.count, clique_hash=first.hash, tree_nodes=first.total_nodes,
                       roots_covered=w.n_roots, roots_total=total_roots,
                       roots_covered_frac=w.n_roots / max(total_roots, 1), kernel_ms=kernel_ms,
                       golden_match=golden, prep_s=round(prep_s, 2),
                       mode="batched+learned: packed search, learned run order and splits, autotuned",
                       learned_model={"label_roots": rep.sample_roots, "loss": rep.loss,
                                      "runs": int(sched.runs.size - 1), "routed_roots": int(sched.eager.sum()),
                                      "chunk_share": args.chunk_share, "route_share": args.route_share,
                                      "seconds": round(learn_s, 2)},
                       tuned=dict(tune, seconds=round(tune_s, 2), stride=args.tune_stride),
                       root_granularity=gran,
                       passes=first.passes, waves=first.waves),
        "plain": plain_info,
        "speedup_vs_plain": value / plain_info["value"],
        "ablation": {"batched_plain_order": {
            "value": abl_outs[0].count * a_steps / abl_wall, "ms_per_step": abl_wall * 1e3 / a_steps,
            "kernel_ms": sum(o.kernel_ms for o in abl_outs) / a_steps, "steps": a_steps},
            "learned_gain": (abl_wall / a_steps) / (wall / args.steps)},
        "roofline": roofline(first, kernel_ms, args.workload, plain),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(g, w.n_roots, args.cpu_seconds, prep_s)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
