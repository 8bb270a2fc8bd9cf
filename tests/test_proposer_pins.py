"""Pins of the proposer oracle (oracle/proposer_oracle.py): the gradient
against central finite differences, the forward pass against closed forms,
and the features against hand-computed values on a small grid.  CPU only."""
import math

import numpy as np

import proposer_oracle as po


def _random_params(rng, scale=0.3):
    return rng.standard_normal(po.NPARAMS) * scale


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _random_params(rng)
    x = rng.standard_normal((40, po.F))
    y = rng.integers(0, 2, 40)
    w = rng.uniform(0.5, 2.0, 40)
    _, g = po.loss_and_grad(p, x, y, w)
    h = 1e-6
    for i in rng.choice(po.NPARAMS, 60, replace=False).tolist() + [po.NPARAMS - 1, 0, po.H * po.F]:
        q = p.copy(); q[i] += h
        lp, _ = po.loss_and_grad(q, x, y, w)
        q[i] -= 2 * h
        lm, _ = po.loss_and_grad(q, x, y, w)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_forward_closed_form():
    # W2 = 0, b2 = c: every second-layer unit is c, so z = c * sum(W3) + b3 for any input
    rng = np.random.default_rng(1)
    p = _random_params(rng)
    W1, b1, W2, b2, W3, b3 = po.unpack(p)
    o = po.H * po.F + po.H
    p[o:o + po.H * po.H] = 0.0
    p[o + po.H * po.H:o + po.H * po.H + po.H] = 0.7
    z, prob, _, _ = po.forward(p, rng.standard_normal((5, po.F)))
    want = 0.7 * np.sum(p[-po.H - 1:-1]) + p[-1]
    assert np.allclose(z, want, rtol=0, atol=1e-12)
    assert np.allclose(prob, 1 / (1 + math.exp(-want)))
    # loss of a perfectly confident correct prediction tends to 0
    q = np.zeros(po.NPARAMS); q[-1] = 40.0
    loss, _ = po.loss_and_grad(q, np.zeros((3, po.F)), [1, 1, 1])
    assert loss < 1e-15


def test_features_by_hand():
    # full 2D grid of max level 2: level/support pairs (0,1), (1,1), (2,1), (2,3) per dimension
    pairs = [(0, 1), (1, 1), (2, 1), (2, 3)]
    level, suppt = [], []
    for a in pairs:
        for b in pairs:
            level.append([a[0], b[0]]); suppt.append([a[1], b[1]])
    E, K = len(level), 4
    u = np.zeros(E * K)
    norms = np.arange(1, E + 1) * 1e-4
    for e in range(E):
        u[e * K] = norms[e]  # one nonzero coefficient: norm = that value
    eps = 1e-3
    feats, leaf = po.features(level, suppt, u, eps, nmax=2)
    f = lambda n: np.clip(np.log10(n / eps) / 4, -4, 4)  # noqa: E731
    # element (l = (2, 2), j = (1, 3)): neighbours in x at j = 3 (wrap both ways), in y at j = 1
    e = level.index([2, 2]) + suppt[level.index([2, 2]):].index([1, 3])
    ex = [i for i in range(E) if level[i] == [2, 2] and suppt[i] == [3, 3]][0]
    ey = [i for i in range(E) if level[i] == [2, 2] and suppt[i] == [1, 1]][0]
    assert np.allclose(feats[e, :5], [f(norms[e]), f(norms[ex]), f(norms[ex]), f(norms[ey]), f(norms[ey])])
    assert np.isclose(feats[e, 5], 4 / 4) and np.isclose(feats[e, 6], 2 / 2)
    assert leaf[e] == 0  # level 2 = nmax in both dimensions: no children possible
    # level (0, 0): its own neighbour, one child per dimension, both present: not a leaf
    e0 = level.index([0, 0])
    assert np.allclose(feats[e0, 1:5], f(norms[e0])) and leaf[e0] == 0
    # drop element (1, 1) x (0, 1): (0,0) loses its x child and becomes a leaf
    keep = [i for i in range(E) if not (level[i] == [1, 0])]
    f2, l2 = po.features([level[i] for i in keep], [suppt[i] for i in keep],
                         np.concatenate([u[i * K:(i + 1) * K] for i in keep]), eps, 2)
    assert l2[keep.index(e0)] == 1
    # a missing same-level neighbour has norm 0: the clamped floor
    keep = [i for i in range(E) if i != ex]
    f3, _ = po.features([level[i] for i in keep], [suppt[i] for i in keep],
                        np.concatenate([u[i * K:(i + 1) * K] for i in keep]), eps, 2)
    assert f3[keep.index(e), 1] == -4.0
