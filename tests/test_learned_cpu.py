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
