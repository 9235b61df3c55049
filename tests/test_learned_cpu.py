"""Learned scheduler, CPU side: the NumPy reference of the MLP gradient is
pinned by central finite differences in float64, and the schedule is a
permutation of the requested roots with a plain-order fallback."""
import numpy as np

import mlp_ref
from rdmce_amd import learned


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.4, mlp_ref.NTRAIN + 2 * mlp_ref.F)
    p[mlp_ref.NTRAIN:mlp_ref.NTRAIN + 8] = rng.normal(0, 1, 8)       # mean
    p[mlp_ref.NTRAIN + 8:] = rng.uniform(0.5, 2.0, 8)                # inv std
    X = rng.normal(0, 1, (40, 8))
    Y = rng.normal(0, 1, 40)
    g = mlp_ref.gradient(p, X, Y)
    eps = 1e-6
    for i in rng.choice(mlp_ref.NTRAIN, 120, replace=False):
        q1, q2 = p.copy(), p.copy()
        q1[i] += eps
        q2[i] -= eps
        fd = (mlp_ref.loss(q1, X, Y) - mlp_ref.loss(q2, X, Y)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), i


def test_learned_schedule_is_a_permutation_with_eager_top():
    rng = np.random.default_rng(1)
    pred = rng.normal(size=1000).astype(np.float32)
    ids = rng.choice(1000, 300, replace=False)
    s = learned.learned_schedule(pred, ids, eager_top=7)
    assert s.learned
    assert sorted(s.order.tolist()) == sorted(ids.tolist())
    assert np.all(np.diff(pred[s.order.astype(np.int64)]) <= 0)
    assert s.eager.sum() == 7 and s.eager[:7].all()


def test_learned_schedule_falls_back_on_garbage():
    pred = np.array([1.0, np.nan, 2.0, 3.0], dtype=np.float32)
    s = learned.learned_schedule(pred, [0, 1, 2, 3], eager_top=2)
    assert not s.learned
    assert s.order.tolist() == [3, 2, 1, 0] and s.eager.sum() == 0


def test_spearman():
    a = np.arange(50)
    assert learned.spearman(a, a * 3 + 1) > 0.999
    assert learned.spearman(a, -a) < -0.999


def test_hybrid_schedule_heavy_first_then_plain():
    """heavy_fraction: the predicted-heaviest share first (longest first),
    the rest in the plain order; always a permutation of the roots."""
    rng = np.random.default_rng(5)
    ids = np.arange(1000, dtype=np.uint64)
    pred = rng.normal(size=1000).astype(np.float32)
    s = learned.learned_schedule(pred, ids, eager_top=3, heavy_fraction=0.05)
    assert s.learned and np.array_equal(np.sort(s.order), ids)
    top = s.order[:50].astype(np.int64)
    assert np.all(np.diff(pred[top]) <= 0)
    assert set(top) == set(np.argsort(-pred, kind="stable")[:50])
    rest = s.order[50:].astype(np.int64)
    assert np.all(np.diff(rest) < 0)  # plain order: decreasing canonical id
    assert s.eager[:3].tolist() == [1, 1, 1] and s.eager[3:].sum() == 0
    full = learned.learned_schedule(pred, ids, heavy_fraction=1.0)
    assert np.all(np.diff(pred[full.order.astype(np.int64)]) <= 0)


def _group_case(seed=0, n_v=60):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 12, size=n_v)
    v0 = np.repeat(np.arange(n_v), sizes)          # canonical roots sorted by first vertex
    ids = np.arange(v0.size, dtype=np.uint64)
    cost = rng.lognormal(2.0, 1.5, size=v0.size)
    heavy = int(np.argmax(sizes))
    cost[v0 == heavy] *= 400.0                      # one heavy vertex with the most roots
    return ids, v0, np.log1p(cost).astype(np.float32), heavy


def test_group_schedule_runs_are_single_vertex_and_cover_all():
    from rdmce_amd import learned
    ids, v0, pred, _ = _group_case()
    sc = learned.learned_group_schedule(pred, ids, v0, waves=16, chunk_share=0.25, route_share=1.0)
    assert sc.learned
    assert sorted(sc.order.tolist()) == ids.tolist()
    r = sc.runs.astype(np.int64)
    assert r[0] == 0 and r[-1] == ids.size and np.all(np.diff(r) > 0)
    for a, b in zip(r[:-1], r[1:]):
        assert np.unique(v0[sc.order[a:b].astype(np.int64)]).size == 1
    # dispatch in decreasing predicted run cost
    c = np.expm1(pred.astype(np.float64))
    run_cost = [c[sc.order[a:b].astype(np.int64)].sum() for a, b in zip(r[:-1], r[1:])]
    assert all(x >= y - 1e-9 * max(x, 1) for x, y in zip(run_cost, run_cost[1:]))


def test_group_schedule_splits_heavy_vertex_and_flags_heavy_roots():
    from rdmce_amd import learned
    ids, v0, pred, heavy = _group_case()
    c = np.expm1(pred.astype(np.float64))
    waves = 16
    share = c.sum() / waves
    sc = learned.learned_group_schedule(pred, ids, v0, waves=waves, chunk_share=0.05, route_share=0.5)
    r = sc.runs.astype(np.int64)
    heavy_runs = sum(1 for a, b in zip(r[:-1], r[1:]) if v0[int(sc.order[a])] == heavy)
    assert heavy_runs > 1                           # the heavy vertex is split over several runs
    want = (c[sc.order.astype(np.int64)] > 0.5 * share).astype(np.uint8)
    assert np.array_equal(sc.eager, want)
    # a chunk's cost exceeds the chunk size by at most one root
    for a, b in zip(r[:-1], r[1:]):
        cc = c[sc.order[a:b].astype(np.int64)]
        assert cc.sum() <= 0.05 * share + cc.max() + 1e-9 * share
    # without splitting, one run per first vertex
    sc1 = learned.learned_group_schedule(pred, ids, v0, waves=waves, chunk_share=1e9, route_share=1e9)
    assert sc1.runs.size - 1 == np.unique(v0).size and sc1.eager.sum() == 0


def test_group_schedule_falls_back_on_nonfinite():
    from rdmce_amd import learned
    ids, v0, pred, _ = _group_case(1)
    pred[3] = np.nan
    sc = learned.learned_group_schedule(pred, ids, v0, waves=8)
    assert not sc.learned and sc.runs is None
    assert np.array_equal(sc.order, ids[::-1]) and sc.eager.sum() == 0
