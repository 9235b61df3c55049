"""Parity cases for the host emulation of the library (test infrastructure).

Run with RDMCE_LIB naming the emulated build (tests/emu/_build/
librdmce_emu.so); every case compares the kernels' results with the CPU
oracle bit for bit (count, histogram, clique hash, per-root node counts and,
where asked, the clique list).  The emulation runs one OS thread per lane,
so the cases are small graphs and few waves (EMU_CUS compute units).

    RDMCE_LIB=tests/emu/_build/librdmce_emu.so EMU_CUS=1 python tests/emu/run_cases.py [filter]
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

import numpy as np  # noqa: E402

import rdmce_amd as R  # noqa: E402
from graphs import closed_forms, dense_core_graph, random_graph  # noqa: E402
from oracle_lib import OracleGraph, clique_hash  # noqa: E402


def oracle_for(g, root_mode=2, root_nodes=True, list_cap=0):
    og = OracleGraph.from_csr(g.GetRowOffset(), g.GetColIdx(), g.GetLabels(), g.GetDegeneracy())
    return og.mce(root_mode=root_mode, root_nodes=root_nodes, list_cap=list_cap)


def same(out, ref, lists=False, nodes=True):
    if (out.count, out.hash, out.hist_dict()) != (ref.count, ref.hash, ref.hist):
        return f"count/hash/hist {out.count} {out.hash} {out.hist_dict()} != {ref.count} {ref.hash} {ref.hist}"
    if nodes and not np.array_equal(out.root_nodes, ref.root_nodes):
        return "per-root node counts differ"
    if lists:
        if sorted(out.cliques) != sorted(tuple(sorted(c)) for c in ref.cliques):
            return "clique lists differ"
        if clique_hash(out.cliques) != out.hash:
            return "clique list does not hash to the reported hash"
    return None


def graph(edges):
    g = R.Graph.from_edges(edges)
    g.SortByOrder(R.OrderType.DEG)
    return g


def cases():
    wpc = 4
    g = R.Graph()
    g.LoadFromFile(os.path.join(TESTS, "golden", "zachary.txt"))
    g.SortByOrder(R.OrderType.DEG)
    yield "zachary list", lambda: same(R.solve(g, waves_per_cu=wpc, want_list=True, want_root_stats=True),
                                       oracle_for(g, list_cap=1 << 16), lists=True)
    for name, edges, _ in closed_forms():
        if name in ("K130", "cocktail20", "moon_moser7"):
            continue
        cg = graph(edges)
        yield f"closed {name}", lambda cg=cg: same(R.solve(cg, waves_per_cu=wpc, want_root_stats=True), oracle_for(cg))
    for n, p, seed in [(60, 0.5, 1), (120, 0.3, 2), (45, 0.8, 3), (200, 0.08, 4)]:
        rg = graph(random_graph(n, p, seed))
        ref = oracle_for(rg)
        for tau in (1, 4, 24, -1):
            for lds in (1, 0):
                yield (f"random G({n},{p}) tau={tau} lds={lds}",
                       lambda rg=rg, ref=ref, tau=tau, lds=lds: same(
                           R.solve(rg, waves_per_cu=wpc, share_bound=tau, lds_tier=lds, want_root_stats=True), ref))
    # bitmap in LDS with the stack in HBM (a small LDS area), and the bitmap
    # tiers alone (word_path off)
    mg = graph(random_graph(150, 0.45, 8))
    mref = oracle_for(mg)
    for lb in (1024, 2048):
        yield (f"mixed tier lds_bytes={lb}",
               lambda lb=lb: same(R.solve(mg, waves_per_cu=wpc, lds_bytes=lb, share_bound=2, want_root_stats=True),
                                  mref))
    yield "word path off", lambda: same(R.solve(mg, waves_per_cu=wpc, word_path=False, want_root_stats=True), mref)
    # dense single-word roots: levels published, taken by thieves and served
    # by their owner
    sg = graph(random_graph(64, 0.85, 11))
    sref = oracle_for(sg)
    yield "word path sharing", lambda: same(R.solve(sg, waves_per_cu=wpc, share_bound=1, want_root_stats=True), sref)
    yield "bitmap tier sharing", lambda: same(R.solve(sg, waves_per_cu=wpc, share_bound=1, word_path=False,
                                                      want_root_stats=True), sref)
    # vertex roots with |P0| in (64, 128] and |X0'| <= 64 (two-word register
    # path) beside single-word ones
    tg = graph(random_graph(300, 0.3, 32))
    tref = oracle_for(tg, root_mode=1)
    for tau in (-1, 2):
        yield (f"two-word vertex-mode roots tau={tau}",
               lambda tau=tau: same(R.solve(tg, waves_per_cu=wpc, root_mode=R.ROOT_VERTEX, share_bound=tau,
                                            want_root_stats=True), tref))
    # the same two-word path, cheaply: three of those vertex roots alone (a
    # custom subset), with and without sharing
    ro, co = tg.GetRowOffset(), tg.GetColIdx()
    outdeg = [int(ro[v + 1]) - int(np.searchsorted(co[ro[v]:ro[v + 1]], v, side="right")) - int(ro[v])
              for v in range(tg.GetNumVertices())]
    sub = np.array([v for v in range(len(outdeg)) if 65 <= outdeg[v] <= 80][:3], dtype=np.uint64)
    ogt = OracleGraph.from_csr(ro, co, tg.GetLabels(), tg.GetDegeneracy())
    sref_sub = ogt.mce_roots(sub, root_mode=1)

    def two_word_subset(tau):
        o = R.solve(tg, waves_per_cu=wpc, root_mode=R.ROOT_VERTEX, share_bound=tau, root_order=sub,
                    want_root_stats=True)
        if len(sub) != 3:
            return "expected three roots with 65..80 candidates"
        if (o.count, o.hash, o.hist_dict()) != (sref_sub.count, sref_sub.hash, sref_sub.hist):
            return "count/hash/hist differ on the two-word subset"
        want = np.zeros_like(tref.root_nodes)
        want[sub.astype(np.int64)] = tref.root_nodes[sub.astype(np.int64)]
        return None if np.array_equal(o.root_nodes, want) else "per-root node counts differ"
    for tau in (-1, 2):
        yield f"two-word subset tau={tau}", lambda tau=tau: two_word_subset(tau)
    vg = graph(random_graph(80, 0.4, 5))
    yield "vertex roots", lambda: same(R.solve(vg, waves_per_cu=wpc, root_mode=R.ROOT_VERTEX, want_root_stats=True),
                                       oracle_for(vg, root_mode=1))
    dg = graph(random_graph(150, 0.35, 6))
    yield "deferred pass", lambda: same(R.solve(dg, waves_per_cu=wpc, slot_bytes=4096, want_root_stats=True),
                                        oracle_for(dg))
    bgs = [graph(random_graph(30 + 10 * i, 0.3 + 0.05 * i, 10 + i)) for i in range(5)]

    def batch():
        b = R.Batch(bgs, waves_per_cu=wpc)
        out = b.run()
        b.free()
        for i, bg in enumerate(bgs):
            ref = oracle_for(bg, root_nodes=False)
            if (int(out.count[i]), int(out.hash[i]), out.hist_dict(i)) != (ref.count, ref.hash, ref.hist):
                return f"batch graph {i} differs"
        return None
    yield "batch", batch
    pg = graph(random_graph(70, 0.5, 7))
    pref = oracle_for(pg, root_nodes=False)

    def replay():
        plan = R.Plan(pg, waves_per_cu=wpc, use_graph=True)
        for _ in range(3):
            o = plan.run()
            if (o.count, o.hash, o.hist_dict()) != (pref.count, pref.hash, pref.hist):
                return "graph replay differs"
        plan.free()
        return None
    yield "plan graph replay", replay

    # packed search (pack_kernel.hip): one tree node per lane, subproblems
    # resident in LDS slots, edge roots derived from shared vertex bitmaps
    yield "pack zachary list", lambda: same(R.solve(g, waves_per_cu=wpc, pack=True, pack_waves_per_cu=2,
                                                    want_list=True, want_root_stats=True),
                                            oracle_for(g, list_cap=1 << 16), lists=True)
    for n, p, seed in [(60, 0.5, 1), (45, 0.8, 3), (150, 0.45, 8)]:
        rg = graph(random_graph(n, p, seed))
        ref = oracle_for(rg)
        for slots in (1, 4):
            yield (f"pack G({n},{p}) slots={slots}",
                   lambda rg=rg, ref=ref, slots=slots: same(
                       R.solve(rg, waves_per_cu=wpc, pack=True, pack_slots=slots, pack_waves_per_cu=2,
                               want_root_stats=True), ref))
    yield "pack vertex two-word", lambda: same(R.solve(tg, waves_per_cu=wpc, root_mode=R.ROOT_VERTEX, pack=True,
                                                       pack_waves_per_cu=2, want_root_stats=True), tref)

    def pack_wide_runs(k):
        # dense vertices whose runs mix one-word, two-word and (k = 150)
        # wider roots, with vertex bitmaps of up to 2 (k = 90) or 3 words
        rg = graph(dense_core_graph(k, 5, 200, 0.05, 11))
        return same(R.solve(rg, waves_per_cu=wpc, pack=True, pack_waves_per_cu=2, want_root_stats=True),
                    oracle_for(rg))
    yield "pack wide runs K90", lambda: pack_wide_runs(90)
    yield "pack wide runs K150", lambda: pack_wide_runs(150)

    def pack_routed_runs():
        # custom order, explicit runs split mid-vertex, every 5th root routed
        # to the one-root-per-wave kernel
        rg = graph(random_graph(120, 0.45, 9))
        ref = oracle_for(rg)
        m = rg.GetNumEdges()
        order = np.arange(m, dtype=np.uint64)[::-1].copy()
        v0 = (rg.roots(R.ROOT_EDGE) >> np.uint64(32)).astype(np.int64)[order.astype(np.int64)]
        cuts = [0] + [i for i in range(1, m) if v0[i] != v0[i - 1] or i % 7 == 0] + [m]
        fl = np.zeros(m, dtype=np.uint8)
        fl[::5] = 1
        out = R.solve(rg, waves_per_cu=wpc, pack=True, pack_waves_per_cu=2, root_order=order, eager=fl,
                      runs=np.array(cuts, dtype=np.uint64), want_root_stats=True)
        return same(out, ref)
    yield "pack routed runs", pack_routed_runs

    def choose_root_mode():
        # vertex roots of every 3rd vertex against those vertices' edge
        # roots: the chooser checks they agree; the oracle checks the count
        rg = graph(random_graph(120, 0.1, 21))
        ref = OracleGraph.from_csr(rg.GetRowOffset(), rg.GetColIdx(), rg.GetLabels(), rg.GetDegeneracy()).mce_roots(
            np.arange(0, rg.GetNumVertices(), 3, dtype=np.uint64), root_mode=1)
        for pack in (False, True):
            ch = R.choose_root_mode(rg, stride=3, pack=pack, waves_per_cu=wpc, pack_waves_per_cu=2)
            if ch["sample_cliques"] != ref.count or ch["root_mode"] not in (R.ROOT_VERTEX, R.ROOT_EDGE):
                return f"chooser pack={pack}: {ch} against {ref.count} cliques"
        return None
    yield "choose root mode", choose_root_mode

    def pack_batch():
        b = R.Batch(bgs, waves_per_cu=wpc, pack=True, pack_waves_per_cu=2)
        out = b.run()
        b.free()
        for i, bg in enumerate(bgs):
            ref = oracle_for(bg, root_nodes=False)
            if (int(out.count[i]), int(out.hash[i]), out.hist_dict(i)) != (ref.count, ref.hash, ref.hist):
                return f"packed batch graph {i} differs"
        return None
    yield "pack batch", pack_batch


def main():
    flt = sys.argv[1] if len(sys.argv) > 1 else ""
    fails = 0
    for name, fn in cases():
        if flt not in name:
            continue
        t = time.time()
        print(f"[emu-case] {name} ...", end=" ", flush=True)
        err = fn()
        print(f"{'ok' if err is None else 'FAIL: ' + err} ({time.time() - t:.1f}s)", flush=True)
        fails += err is not None
    print(f"[emu] {fails} failed", flush=True)
    sys.exit(1 if fails else 0)


if __name__ == "__main__":
    main()
