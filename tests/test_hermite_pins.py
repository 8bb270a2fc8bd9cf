"""Pins of the oracle's Hermite P3 flux interpolation (HermBasis,
source/HermBasis.cpp:9-250,347-552; HermInterpolation, source/Interplation.cpp:
1886-2316,2921-3030,3166-3316,3651-3850; HyperbolicHermRHS /
HyperbolicSameFluxHermRHS, source/FastMultiplyLU.cpp:970-1016,1221-1267).  CPU only.

The reference ships no test of this path, so each piece is checked against
what the mathematics fixes:
  * every basis function takes value or slope 1 at its own degree of freedom
    and 0 at the others of its cell (all levels: pins the level scaling);
  * the slope functions are the derivatives of the value functions (finite
    differences);
  * point values / slopes of a cubic solution are its closed-form derivatives
    (pins the Alpert-at-Hermite-point table);
  * the point flux is the chain rule of f(u) (pins the slope and cross-slope
    formulas for Burgers and sin/cos);
  * hierarchical coefficients of a cubic vanish above level 0 and the
    interpolant reproduces it (pins the surplus weights);
  * the interpolation error of a smooth flux falls like h^4 (all levels);
  * the right-hand side from the Hermite family equals the Lagrange family's
    when both interpolate the flux exactly (pins the Hermite-Alpert tables);
  * whole runs with either family approach each other as the grid is refined.
"""
import math

import numpy as np
import pytest

import pyoracle
from common import oracle_run_hyperbolic, problem

PMAX = 3


def order(level, suppt, p):
    slot = 0 if level == 0 else (1 << (level - 1)) + (suppt - 1) // 2
    return slot * (PMAX + 1) + p


def hermite_oracle(dim, k, nmax, sparse, level_init):
    o = pyoracle.Oracle(dim, k, 3, nmax, sparse, level_init)
    o.set_interpolation(1)
    return o


CELLS = [(0, 1), (1, 1), (2, 1), (2, 3), (3, 5), (4, 9), (5, 31)]


def test_basis_interpolates_its_degrees_of_freedom():
    o = hermite_oracle(1, 2, 5, False, 0)
    for level, suppt in CELLS:
        for p in range(4):
            b = order(level, suppt, p)
            for q in range(4):
                x = o.lagr_point(order(level, suppt, q))
                v = o.basis_val(2, b, x, q // 2)
                # the points sit ROUND_OFF * h / 2 inside the half cells: slopes of the
                # value functions there are O(2^level * ROUND_OFF)
                assert abs(v - (1.0 if p == q else 0.0)) < 1e-12 * 2.0 ** (level + 2), (level, suppt, p, q, v)


def test_points_sit_at_cell_ends_and_midpoints():
    o = hermite_oracle(1, 2, 5, False, 0)
    for level, suppt in CELLS:
        if level == 0:
            want = [0, 1, 0, 1]
        else:
            h = 2.0 ** (-(level - 1))
            mid = h * (suppt - 1) / 2 + h / 2
            want = [mid] * 4
        got = [o.lagr_point(order(level, suppt, q)) for q in range(4)]
        assert np.allclose(got, want, atol=1e-12), (level, suppt, got)
        # value and slope of one point share its position exactly
        assert got[0] == got[2] and got[1] == got[3]


def test_slope_functions_are_derivatives_of_the_values():
    o = hermite_oracle(1, 2, 5, False, 0)
    rng = np.random.default_rng(4)
    h = 1e-6
    for level, suppt in CELLS:
        lo = 0.0 if level <= 1 else 2.0 ** (-(level - 1)) * (suppt - 1) / 2
        w = 1.0 if level <= 1 else 2.0 ** (-(level - 1))
        for p in range(4):
            b = order(level, suppt, p)
            for t in rng.uniform(0.02, 0.98, 6):
                if abs(t - 0.5) < 0.02:
                    continue  # the level >= 1 functions end at the cell midpoint
                x = lo + w * t
                fd = (o.basis_val(2, b, x + h * w) - o.basis_val(2, b, x - h * w)) / (2 * h * w)
                d = o.basis_val(2, b, x, 1)
                assert abs(fd - d) < 1e-6 * max(1.0, abs(d)), (level, suppt, p, x, fd, d)


# a separable cubic solution u = g(x) h(y), exactly representable with k = 3
G = np.polynomial.Polynomial([1.0, 1.0, -0.7, 0.4])
H = np.polynomial.Polynomial([0.5, 0.0, -1.0, 0.2])


def _cubic_solution(sparse, nmax, g=G, h=H):
    o = hermite_oracle(2, 3, nmax, sparse, nmax)
    o.init_separable([lambda x, d: float(g(x)) if d == 0 else float(h(x))], 1e10)
    return o


def _dof_points(o):
    """For every element and degree of freedom (p0, p1): the point and derivative orders."""
    lev, sup, _ = o.grid()
    n = len(lev)
    x = np.zeros((n, 16, 2))
    der = np.zeros((n, 16, 2), dtype=int)
    for e in range(n):
        for q in range(16):
            pq = (q // 4, q % 4)
            for d in range(2):
                x[e, q, d] = o.lagr_point(order(lev[e, d], sup[e, d], pq[d]))
                der[e, q, d] = pq[d] // 2
    return x, der


def test_point_values_and_slopes_of_a_cubic():
    o = _cubic_solution(False, 3)
    o.eval_up()
    up = o.get(pyoracle.FIELD_UP).reshape(-1, 16)
    x, der = _dof_points(o)
    gd = [G, G.deriv()]
    hd = [H, H.deriv()]
    want = np.zeros_like(up)
    for e in range(up.shape[0]):
        for q in range(16):
            want[e, q] = gd[der[e, q, 0]](x[e, q, 0]) * hd[der[e, q, 1]](x[e, q, 1])
    assert np.max(np.abs(up - want)) < 1e-11


@pytest.mark.parametrize("kind", [0, 1])
def test_point_flux_is_the_chain_rule(kind):
    o = _cubic_solution(True, 4)
    o.eval_up()
    o.flux(kind, (1.0,))
    x, der = _dof_points(o)
    u = G(x[..., 0]) * H(x[..., 1])
    ux = G.deriv()(x[..., 0]) * H(x[..., 1])
    uy = G(x[..., 0]) * H.deriv()(x[..., 1])
    uxy = G.deriv()(x[..., 0]) * H.deriv()(x[..., 1])
    if kind == 0:
        comps = [(u * u / 2, u, np.ones_like(u))] * 2
    else:
        comps = [(np.sin(u), np.cos(u), -np.sin(u)), (np.cos(u), -np.sin(u), -np.cos(u))]
    for c, (f, f1, f2) in enumerate(comps):
        fx, fy, fxy = f1 * ux, f1 * uy, f2 * ux * uy + f1 * uxy
        want = np.where((der[..., 0] == 0) & (der[..., 1] == 0), f,
                        np.where(der[..., 1] == 0, fx, np.where(der[..., 0] == 0, fy, fxy)))
        got = o.get(pyoracle.FIELD_FP + c).reshape(-1, 16)
        assert np.max(np.abs(got - want)) < 1e-10, (kind, c)


def _hermite_table_1d(o, nmax, xs):
    """[slot*4 + p, i]: Hermite basis function (slot, p) at xs[i]."""
    n = (1 << nmax) * 4
    t = np.zeros((n, len(xs)))
    for r in range(n):
        for i, xv in enumerate(xs):
            t[r, i] = o.basis_val(2, r, xv)
    return t


def _interpolant(o, nmax, comp, xs, ys):
    lev, sup, _ = o.grid()
    coe = o.get(pyoracle.FIELD_FUCOE + comp).reshape(-1, 4, 4)
    tx = _hermite_table_1d(o, nmax, xs)
    ty = tx if ys is xs else _hermite_table_1d(o, nmax, ys)
    val = np.zeros((len(xs), len(ys)))
    for e in range(len(lev)):
        s0 = order(lev[e, 0], sup[e, 0], 0) // 4
        s1 = order(lev[e, 1], sup[e, 1], 0) // 4
        val += tx[s0 * 4:s0 * 4 + 4].T @ coe[e] @ ty[s1 * 4:s1 * 4 + 4]
    return val


@pytest.mark.parametrize("sparse", [False, True])
def test_surpluses_of_a_cubic_vanish_and_the_interpolant_is_exact(sparse):
    # u linear in each variable: the Burgers flux u^2/2 is quadratic, inside the P3 space
    g = np.polynomial.Polynomial([0.4, 0.9])
    h = np.polynomial.Polynomial([1.1, -0.6])
    nmax = 4
    o = _cubic_solution(sparse, nmax, g, h)
    o.eval_up()
    o.flux(0, (0.0,))
    o.fp_to_coe()
    lev, _, _ = o.grid()
    coe = o.get(pyoracle.FIELD_FUCOE).reshape(len(lev), 16)
    assert np.max(np.abs(coe[lev.sum(axis=1) > 0])) < 1e-11
    xs = np.random.default_rng(8).uniform(0, 1, 9)
    got = _interpolant(o, nmax, 0, xs, xs)
    u = np.outer(g(xs), h(xs))
    assert np.max(np.abs(got - u * u / 2)) < 1e-11


def test_interpolation_error_falls_like_h4():
    # f = sin(u) of the cubic u: smooth, so the P3 Hermite interpolation error is O(h^4)
    xs = np.random.default_rng(3).uniform(0.01, 0.99, 7)
    u = np.outer(G(xs), H(xs))
    errs = []
    for nmax in (2, 3, 4):
        o = _cubic_solution(False, nmax)
        o.eval_up()
        o.flux(1, (1.0,))
        o.fp_to_coe()
        errs.append(np.max(np.abs(_interpolant(o, nmax, 0, xs, xs) - np.sin(u))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 3.3 and errs[-1] < 1e-5, (errs, rates)


@pytest.mark.parametrize("sparse", [False, True])
def test_rhs_matches_the_lagrange_family_on_exact_fluxes(sparse):
    g = np.polynomial.Polynomial([0.4, 0.9])
    h = np.polynomial.Polynomial([1.1, -0.6])
    out = []
    for fam in (0, 1):
        o = pyoracle.Oracle(2, 2, 3, 4, sparse, 4)
        o.init_separable([lambda x, d: float(g(x)) if d == 0 else float(h(x))], 1e10)
        o.set_interpolation(fam)
        o.eval_up()
        o.flux(0, (0.0,))
        o.fp_to_coe()
        o.rhs_zero()
        o.rhs_lagr_vol()
        o.rhs_lagr_flx()
        out.append(o.get(pyoracle.FIELD_RHS))
    assert np.max(np.abs(out[0])) > 1e-2
    assert np.max(np.abs(out[0] - out[1])) < 1e-11


def test_one_dimensional_flux_and_family_switch():
    # 1D: the slope dof of f(u) is f'(u) u_x; switching back restores the Lagrange result
    g = np.polynomial.Polynomial([0.3, -0.8, 0.5, 0.25])
    o = pyoracle.Oracle(1, 3, 3, 4, False, 4)
    o.init_separable([lambda x, d: float(g(x))], 1e10)
    o.eval_up()
    lag = o.get(pyoracle.FIELD_UP)
    o.set_interpolation(1)
    o.eval_up()
    o.flux(0, (0.0,))
    lev, sup, _ = o.grid()
    fp = o.get(pyoracle.FIELD_FP).reshape(-1, 4)
    for e in range(len(lev)):
        for q in range(4):
            x = o.lagr_point(order(lev[e, 0], sup[e, 0], q))
            want = g(x) ** 2 / 2 if q < 2 else g(x) * g.deriv()(x)
            assert abs(fp[e, q] - want) < 1e-11
    o.set_interpolation(0)
    o.eval_up()
    assert np.array_equal(o.get(pyoracle.FIELD_UP), lag)


@pytest.mark.parametrize("kind", [0, 1])
def test_hermite_and_lagrange_runs_approach_each_other(kind):
    # smooth data, 0.05 time units: the two flux interpolations differ by O(h^3) or better
    diffs = []
    for n in (3, 4):
        out = []
        for fam in (0, 1):
            o = pyoracle.Oracle(2, 2, 3, n, False, n)
            o.init_separable([lambda x, d: 0.5 if d == 0 else 1.0,
                              lambda x, d: 0.25 * math.sin(2 * math.pi * x) if d == 0 else math.cos(2 * math.pi * x)],
                             1e10)
            o.set_interpolation(fam)
            prob = problem(flux_kind=kind, flux_param=1.0, alpha=[1.0, 1.0], wave_speed=[1.0, 1.0], cfl=0.1,
                           t_final=0.05, adaptive=0, eps=1e-3, eta=1e-4, max_steps=10 ** 6)
            oracle_run_hyperbolic(o, prob)
            out.append(o.get(pyoracle.FIELD_UCOE))
        diffs.append(np.sqrt(np.sum((out[0] - out[1]) ** 2)))
    assert diffs[1] < diffs[0] / 6 and diffs[1] < 3e-5, diffs
