"""Pins of the oracle's separable initialisation (DGAdapt::init_separable_scalar_sum,
source/DGAdapt.cpp:120-147,850-894), DGAdapt::filter / filter_local
(source/DGAdapt.cpp:461-555), point values (DGSolution::val,
source/DGSolution.cpp:613-625) and the quadrature error
(get_error_no_separable_scalar, source/DGSolution.cpp:326-335).  CPU only.

Independent checks: the 1D projections against numpy Gauss-Legendre
quadrature of high order on fine cells, polynomial reproduction, the
refinement fixed point (no leaf above eps, no holes), linearity in the
functions, Parseval for sin(2 pi (x + y)), filter_local against filter for a
constant speed field, point values of projected polynomials, and the
quadrature error of an orthogonal projection against Pythagoras."""
import math

import numpy as np
import pytest

import pyoracle

PI = math.pi


def _order(level, suppt, p, k):
    slot = 0 if level == 0 else (1 << (level - 1)) + (suppt - 1) // 2
    return slot * (k + 1) + p


def _support(level, suppt):
    if level <= 1:
        return 0.0, 1.0
    h = 2.0 ** (-level + 1)
    return h * (suppt - 1) / 2, h * (suppt + 1) / 2


def _projection(o, f, level, suppt, p, k):
    """<f, v> over the support by 24-point Gauss-Legendre on 32 cells (numpy)."""
    xg, wg = np.polynomial.legendre.leggauss(24)
    a, b = _support(level, suppt)
    cells = np.linspace(a, b, 33)
    order = _order(level, suppt, p, k)
    tot = 0.0
    for i in range(32):
        lo, hi = cells[i], cells[i + 1]
        x = (xg + 1) / 2 * (hi - lo) + lo
        v = np.array([o.basis_val(0, order, xi) for xi in x])
        tot += np.sum(wg * v * f(x)) * (hi - lo) / 2
    return tot


def test_projection_matches_independent_quadrature():
    k, nmax = 2, 5
    o = pyoracle.Oracle(2, k, 3, nmax, True, nmax)
    fx = [lambda x: np.sin(2 * PI * x) + 0.3 * x, lambda x: np.exp(np.cos(2 * PI * x))]
    o.init_separable([lambda x, d: float(fx[d](np.array([x]))[0])], 1e10)
    lev, sup, _ = o.grid()
    u = o.get(pyoracle.FIELD_UCOE).reshape(-1, k + 1, k + 1)
    rng = np.random.default_rng(5)
    for e in rng.choice(len(lev), 25, replace=False):
        c = [[_projection(o, fx[d], lev[e, d], sup[e, d], p, k) for p in range(k + 1)] for d in range(2)]
        want = np.outer(c[0], c[1])
        assert np.max(np.abs(u[e] - want)) < 1e-12, (e, lev[e], u[e], want)


def test_polynomials_project_onto_level_zero_and_do_not_refine():
    k = 2
    o = pyoracle.Oracle(2, k, 3, 6, True, 2)
    n0 = o.num_elem
    o.init_separable([lambda x, d: 1.0 + x - 0.7 * x * x, lambda x, d: 0.5 - x * x * (d + 1)], 1e-10)
    assert o.num_elem == n0
    lev, _, _ = o.grid()
    u = o.get(pyoracle.FIELD_UCOE).reshape(len(lev), -1)
    assert np.max(np.abs(u[lev.sum(axis=1) > 0])) < 1e-14


def test_refinement_reaches_a_fixed_point_without_holes():
    eps = 1e-3
    o = pyoracle.Oracle(2, 2, 3, 7, True, 1)
    o.init_separable([lambda x, d: math.exp(-((x - 0.4) / 0.08) ** 2)], eps)
    lev, sup, key = o.grid()
    u = o.get(pyoracle.FIELD_UCOE).reshape(len(lev), -1)
    norm = np.sqrt(np.sum(u * u, axis=1))
    present = {(tuple(l), tuple(s)) for l, s in zip(lev.tolist(), sup.tolist())}
    nmax = 7

    def children(l, s):
        out = []
        for d in range(2):
            if l[d] == nmax:
                continue
            l2 = list(l)
            l2[d] += 1
            js = [s[d]] if l[d] == 0 else [2 * s[d] - 1, 2 * s[d] + 1]
            for j in js:
                s2 = list(s)
                s2[d] = j
                out.append((tuple(l2), tuple(s2)))
        return out

    def parents(l, s):
        out = []
        for d in range(2):
            if l[d] == 0:
                continue
            l2, s2 = list(l), list(s)
            l2[d] -= 1
            s2[d] = 1 if l2[d] == 0 else (s[d] + 1) // 2 - ((s[d] + 1) // 2 + 1) % 2
            out.append((tuple(l2), tuple(s2)))
        return out

    refined = 0
    for i, (l, s) in enumerate(zip(lev.tolist(), sup.tolist())):
        ch = children(l, s)
        if any(c not in present for c in ch):  # a leaf
            assert norm[i] <= eps, (l, s, norm[i])
        for p in parents(l, s):
            assert p in present, ("hole", l, s, p)
        refined += len(ch) > 0 and all(c in present for c in ch)
    assert o.num_elem > 20 and refined > 0


def test_linear_in_the_functions():
    f1 = lambda x, d: math.sin(2 * PI * x) if d == 0 else math.cos(2 * PI * x)  # noqa: E731
    f2 = lambda x, d: math.cos(2 * PI * x) if d == 0 else math.sin(2 * PI * x)  # noqa: E731
    a = pyoracle.Oracle(2, 2, 3, 4, True, 4)
    b = pyoracle.Oracle(2, 2, 3, 4, True, 4)
    c = pyoracle.Oracle(2, 2, 3, 4, True, 4)
    a.init_separable([f1, f2], 1e10)
    b.init_separable([f1], 1e10)
    c.init_separable([f2], 1e10)
    assert np.array_equal(a.get(pyoracle.FIELD_UCOE), b.get(pyoracle.FIELD_UCOE) + c.get(pyoracle.FIELD_UCOE))


def test_parseval_for_sin_sum():
    # ||sin(2 pi (x + y))||^2 on the unit square is 1/2; orthonormal coefficients sum to it
    errs = []
    for n in (4, 6):
        o = pyoracle.Oracle(2, 2, 3, n, False, n)
        o.init_separable([lambda x, d: math.sin(2 * PI * x) if d == 0 else math.cos(2 * PI * x),
                          lambda x, d: math.cos(2 * PI * x) if d == 0 else math.sin(2 * PI * x)], 1e10)
        u = o.get(pyoracle.FIELD_UCOE)
        errs.append(abs(np.sum(u * u) - 0.5))
    assert errs[1] < errs[0] and errs[1] < 1e-9, errs


@pytest.mark.parametrize("dim", [2, 3])
def test_filter_local_with_constant_speed_equals_filter(dim):
    rng = np.random.default_rng(11)
    a = pyoracle.Oracle(dim, 1, 2, 4, True, 4)
    b = pyoracle.Oracle(dim, 1, 2, 4, True, 4)
    u = rng.standard_normal(a.num_elem * 2 ** dim)
    a.set(pyoracle.FIELD_UCOE, u)
    b.set(pyoracle.FIELD_UCOE, u)
    speed = [0.7, -1.3, 0.4][:dim]
    a.filter(0.8, [abs(s) for s in speed], 1e-2, 3)
    b.filter_local(0.8, lambda x, d: speed[d], 1e-2, 3)
    assert np.array_equal(a.get(pyoracle.FIELD_UCOE), b.get(pyoracle.FIELD_UCOE))


def test_filter_damps_fine_levels_only():
    rng = np.random.default_rng(2)
    o = pyoracle.Oracle(2, 2, 3, 5, True, 5)
    lev, _, _ = o.grid()
    u = rng.standard_normal(o.num_elem * 9)
    o.set(pyoracle.FIELD_UCOE, u)
    o.filter(1.0, [1.0, 2.0], 1e-2, 4)
    v = o.get(pyoracle.FIELD_UCOE).reshape(-1, 9)
    u = u.reshape(-1, 9)
    ls = lev.sum(axis=1)
    assert np.array_equal(v[ls < 4], u[ls < 4])
    ratio = v[ls >= 4] / u[ls >= 4]
    assert np.all(ratio > 0) and np.all(ratio < 1)
    assert np.allclose(ratio, ratio[:, :1], rtol=0, atol=1e-15)  # one factor per element
    # the factor falls with sum_d a_d 2^l_d
    s = (1.0 * 2.0 ** lev[ls >= 4, 0] + 2.0 * 2.0 ** lev[ls >= 4, 1])
    r = ratio[:, 0]
    order = np.argsort(s)
    assert np.all(np.diff(r[order]) <= 1e-15)


def test_point_values_reproduce_projected_polynomials():
    # a separable polynomial of degree <= k is its own projection: u_h(x) = f(x) everywhere,
    # at interior points and at cell boundaries (closed supports, several cells)
    o = pyoracle.Oracle(2, 2, 3, 4, True, 4)
    f0 = lambda x: 1.0 + x - 0.7 * x * x  # noqa: E731
    f1 = lambda x: 0.5 - 2 * x * x  # noqa: E731
    o.init_separable([lambda x, d: f0(x) if d == 0 else f1(x)], 1e10)
    x1d = np.array([0.0, 0.1, 0.25, 0.3, 0.5, 0.77, 0.875, 1.0])
    v = o.eval_tensor(x1d)
    want = np.outer(f0(x1d), f1(x1d))
    assert np.max(np.abs(v - want)) < 1e-13
    err = o.error_no_separable(lambda x: f0(x[0]) * f1(x[1]), 3)
    assert np.all(err < 1e-13), err


def test_quadrature_error_is_pythagoras_for_a_projection():
    # u_h = orthogonal projection of f onto the full grid space: ||f - u_h||^2 = ||f||^2 - sum c^2
    o = pyoracle.Oracle(2, 1, 2, 3, False, 3)
    o.init_separable([lambda x, d: math.sin(2 * PI * x)], 1e10)
    u = o.get(pyoracle.FIELD_UCOE)
    err = o.error_no_separable(lambda x: math.sin(2 * PI * x[0]) * math.sin(2 * PI * x[1]), 8)
    pyth = math.sqrt(0.25 - np.sum(u * u))
    assert abs(err[1] - pyth) < 1e-9 * pyth, (err[1], pyth)
    assert err[0] <= err[1] <= err[2]
