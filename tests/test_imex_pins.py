"""Pins of the oracle's assembled artificial viscosity and IMEX43
(ArtificialViscosity::assemble_matrix, source/BilinearForm.cpp:1190-1349;
Basis::product_volume_interv, source/Basis.cpp:83-119; IMEX43,
source/ODESolver.cpp:872-1070, with Eigen's conjugate gradient): 1D integrals
against independent quadrature, the quadratic form of a polynomial against its
closed-form Dirichlet integral, the related-element sets against brute force,
IMEX43 without viscosity against SSP-RK3, its temporal order, and the implicit
solve against a dense solve.  CPU only."""
import math

import numpy as np
import pytest

import pyoracle
from test_setup_pins import fine_integral, order

PI = math.pi
G = np.polynomial.Polynomial([0.3, -1.0, 0.8])
H = np.polynomial.Polynomial([1.0, 0.4, -0.6])


def support(level, suppt):
    if level <= 1:
        return 0.0, 1.0
    h = 2.0 ** (1 - level)
    return h * (suppt - 1) / 2, h * (suppt + 1) / 2


@pytest.mark.parametrize("k", [1, 2, 3])
def test_product_volume_interv_independent_quadrature(k):
    nmax = 4
    o = pyoracle.Oracle(1, k, 3, nmax, False, 0)
    rng = np.random.default_rng(k)
    cases = 0
    for _ in range(300):
        lv = rng.integers(0, nmax + 1, 3)
        sp = [1 if l <= 1 else 2 * int(rng.integers(0, 2 ** (l - 1))) + 1 for l in lv]
        p = rng.integers(0, k + 1, 3)
        ru, rv, rw = (order(int(lv[i]), sp[i], int(p[i]), k) for i in range(3))
        a, b = support(int(lv[2]), sp[2])
        for du in (0, 1):
            got = o.product_volume_interv(ru, rv, rw, du, du, k + 1 if du == 0 else k)
            f = lambda x: o.basis_val(0, ru, x, du) * o.basis_val(0, rv, x, du)  # noqa: E731
            want = fine_integral(f, a, b, 16)
            assert abs(got - want) < 1e-10 * max(1.0, abs(want)), (lv, sp, p, du, got, want)
            cases += abs(want) > 1e-8
    assert cases > 50


def grid_with_visc(option, coef=-0.7, nmax=4, dim=2):
    o = pyoracle.Oracle(dim, 2, 3, nmax, True, nmax)
    o.init_separable([lambda x, d: float(G(x)) if d == 0 else float(H(x))], 1e10)
    n, flags = o.viscosity_elements(-100.0)  # every childless element at the finest level
    assert n > 0
    nnz = o.viscosity_operator(coef, option)
    rows, cols, vals = o.viscosity_operator_coo()
    assert nnz == len(vals)
    N = o.num_elem * 9
    A = np.zeros((N, N))
    A[rows, cols] = vals
    return o, flags, A


@pytest.mark.parametrize("option", [1, 2])
def test_quadratic_form_is_the_dirichlet_integral(option):
    # u = G(x) H(y) lives on the level-0 element, which every related set holds, so
    # u^T A u = coef * sum_w int_{supp w} |grad u|^2
    coef = -0.7
    o, flags, A = grid_with_visc(option, coef)
    u = o.get(pyoracle.FIELD_UCOE)
    lv, sp, _ = o.grid()
    want = 0.0
    Gi, Hi = (G * G).integ(), (H * H).integ()
    Gd, Hd = (G.deriv() * G.deriv()).integ(), (H.deriv() * H.deriv()).integ()
    for e in np.nonzero(flags)[0]:
        x0, x1 = support(lv[e, 0], sp[e, 0])
        y0, y1 = support(lv[e, 1], sp[e, 1])
        want += coef * ((Gd(x1) - Gd(x0)) * (Hi(y1) - Hi(y0)) + (Gi(x1) - Gi(x0)) * (Hd(y1) - Hd(y0)))
    got = u @ A @ u
    assert abs(got - want) < 1e-10 * abs(want), (got, want)


@pytest.mark.parametrize("option", [1, 2])
def test_operator_structure(option):
    o, flags, A = grid_with_visc(option)
    assert np.array_equal(A, A.T)  # the 1D integrals are symmetric in (u, v): exact
    ev = np.linalg.eigvalsh(A)
    assert ev.max() < 1e-14 * -ev.min()  # -nu * stiffness: negative semidefinite (to rounding)
    const = np.zeros(A.shape[0])
    lv, sp, _ = o.grid()
    e0 = [e for e in range(o.num_elem) if not lv[e].any()][0]
    const[e0 * 9] = 1.0
    assert not (A @ const).any()  # constants have no gradient: those columns hold no entry
    # related sets against brute force on the grid
    K = 9
    visc = np.nonzero(flags)[0]

    def contains(a, b):  # support of element a holds that of b in every dimension, a no finer
        for d in range(2):
            if lv[a, d] > lv[b, d]:
                return False
            a0, a1 = support(lv[a, d], sp[a, d])
            b0, b1 = support(lv[b, d], sp[b, d])
            if not (a0 <= b0 and b1 <= a1):
                return False
        return True

    def meets(a, b):
        for d in range(2):
            a0, a1 = support(lv[a, d], sp[a, d])
            b0, b1 = support(lv[b, d], sp[b, d])
            if a0 >= b1 or a1 <= b0:
                return False
        return True

    rel = [{e for e in range(o.num_elem) if (contains(e, w) if option == 2 else meets(e, w))} for w in visc]
    allowed = set()
    for s in rel:
        allowed |= {(a, b) for a in s for b in s}
    blocks = {(r // K, c // K) for r, c in zip(*np.nonzero(A))}
    assert blocks <= allowed
    # a viscosity element couples to itself, and the coupled blocks span more than the diagonal
    for w in visc:
        assert A[w * K:(w + 1) * K, w * K:(w + 1) * K].any()
    assert len(blocks) > len(visc)
    if option == 1:
        _, _, A2 = grid_with_visc(2)
        p1, p2 = set(zip(*np.nonzero(A))), set(zip(*np.nonzero(A2)))
        assert p2 < p1  # choice 2 is a subset of choice 1, strictly on this grid


def test_rhs_viscosity_operator_is_the_product():
    o, _, A = grid_with_visc(1)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(A.shape[0])
    r0 = rng.standard_normal(A.shape[0])
    o.set(pyoracle.FIELD_UCOE, u)
    o.set(pyoracle.FIELD_RHS, r0)
    o.rhs_viscosity_operator()
    got = o.get(pyoracle.FIELD_RHS)
    assert np.max(np.abs(got - (r0 + A @ u))) < 1e-12 * np.max(np.abs(A @ u))


def _imex_run(o, dt, steps, coef_lin=(1.0, 1.0)):
    iters = []
    for _ in range(steps):
        o.imex43_init(dt)
        for st in range(5):
            if st >= 2:
                o.rhs_zero()
                o.rhs_linear(list(coef_lin))
            iters.append(o.imex43_stage(st))
    return iters


def _advection(n=3):
    o = pyoracle.Oracle(2, 2, 3, n, True, n)
    o.init_adaptive_intp(lambda x: math.sin(2 * PI * (x[0] + x[1])), 1e10)
    return o


def test_imex43_without_viscosity_is_ssp_rk3():
    # with G = 0 the stages are u2 = un, u3 = un + dt F(un), u4 = un + dt/4 (F(u2) + F(u3)) and the
    # final combination is SSP-RK3 in Butcher form
    dt, steps = 0.004, 10
    a = _advection()
    a.viscosity_elements(1e9)  # no viscosity element: the operator is empty
    assert a.viscosity_operator(-1.0, 1) == 0
    iters = _imex_run(a, dt, steps)
    assert set(iters) == {0}
    b = _advection()
    for _ in range(steps):
        b.rk_init()
        for st in range(3):
            b.rhs_zero()
            b.rhs_linear([1.0, 1.0])
            b.rk_stage(3, st, dt)
    ua, ub = a.get(pyoracle.FIELD_UCOE), b.get(pyoracle.FIELD_UCOE)
    assert np.max(np.abs(ua - ub)) < 1e-13 * np.max(np.abs(ub))


def test_imex43_implicit_solve_and_temporal_order():
    sols = []
    T = 0.04
    for nsteps in (10, 20, 40):
        o = _advection()
        o.viscosity_elements(-100.0)
        o.viscosity_operator(-0.05, 1)
        if nsteps == 10:
            rows, cols, vals = o.viscosity_operator_coo()
            N = o.num_elem * 9
            A = np.zeros((N, N))
            A[rows, cols] = vals
            dt = T / nsteps
            un = o.get(pyoracle.FIELD_UCOE)
            o.imex43_init(dt)
            it = o.imex43_stage(0)
            assert it > 0
            want = np.linalg.solve(np.eye(N) - dt * 0.24169426078821 * A, un)
            assert np.max(np.abs(o.get(pyoracle.FIELD_UCOE) - want)) < 1e-10 * np.max(np.abs(want))
            o.set(pyoracle.FIELD_UCOE, un)
        iters = _imex_run(o, T / nsteps, nsteps)
        assert min(iters) >= 0
        sols.append(o.get(pyoracle.FIELD_UCOE))
    r = np.linalg.norm(sols[0] - sols[1]) / np.linalg.norm(sols[1] - sols[2])
    assert abs(math.log2(r) - 3) < 0.3, (r, math.log2(r))


def test_imex43_is_stable_where_explicit_viscosity_is_not():
    # a large viscosity: the explicit SSP-RK3 with rhs += A u grows without bound at this step,
    # IMEX43 (implicit in A) stays bounded
    dt, steps, coef = 0.01, 30, -2.0
    a = _advection()
    a.viscosity_elements(-100.0)
    a.viscosity_operator(coef, 1)
    _imex_run(a, dt, steps)
    b = _advection()
    b.viscosity_elements(-100.0)
    b.viscosity_operator(coef, 1)
    for _ in range(steps):
        b.rk_init()
        for st in range(3):
            b.rhs_zero()
            b.rhs_linear([1.0, 1.0])
            b.rhs_viscosity_operator()
            b.rk_stage(3, st, dt)
    na, nb = np.max(np.abs(a.get(pyoracle.FIELD_UCOE))), np.max(np.abs(b.get(pyoracle.FIELD_UCOE)))
    assert na < 2.0 and nb > 1e3, (na, nb)
