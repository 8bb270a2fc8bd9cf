"""Pins of the host setup (bases, quadrature, 1D tables, hash key) and of the
oracle's transforms, by the reference's own unit tests, closed forms and
brute force.  CPU only."""
import itertools
import math

import numpy as np
import pytest

import pyoracle

PI = math.pi


def order(level, suppt, p, pmax):
    slot = 0 if level == 0 else (1 << (level - 1)) + (suppt - 1) // 2
    return slot * (pmax + 1) + p


# --- the reference's gtests, restated -------------------------------------------------------------

def test_alpert_orthonormal_reference_case():
    # test/test_AlptBasis.cpp:8-36: (1,1,2) vs (2,3,1), k=2, 10-point GL, TOL 1e-14
    o = pyoracle.Oracle(1, 2, 3, 3, False, 0)
    a, b = order(1, 1, 2, 2), order(2, 3, 1, 2)
    assert abs(o.product_volume(0, a, 0, b)) < 1e-14
    assert abs(o.product_volume(0, a, 0, a) - 1.0) < 1e-14
    assert abs(o.product_volume(0, b, 0, b) - 1.0) < 1e-14


@pytest.mark.parametrize("n", range(1, 11))
def test_gauss_legendre_exactness(n):
    # test/test_Quad.cpp:15-60: n points integrate x^i exactly for i <= 2n-1 on [-1,1]
    x, w = pyoracle.gauss_legendre(n)
    for i in range(2 * n):
        exact = 0.0 if i % 2 else 2.0 / (i + 1)
        assert abs(np.sum(w * x ** i) - exact) < 5e-15 * 2 * 4


def test_gauss_legendre_multid_odd():
    # test/test_Quad.cpp:100-144: odd polynomials integrate to zero up to dim 4
    x, w = pyoracle.gauss_legendre(5)
    for dim in range(1, 5):
        tot = 0.0
        for idx in itertools.product(range(5), repeat=dim):
            pt = x[list(idx)]
            tot += np.prod(w[list(idx)]) * pt[0] ** 3 * np.prod(pt[1:] ** 2)
        assert abs(tot) < 1e-14


def test_tensor_layout_row_major():
    # test/test_VecMultiD.cpp:158-178: iterator order {0,0},{0,1},{0,2},{1,0},... (last index
    # fastest).  The surplus of the level-0 element is the point value itself; with
    # f(x, y) = x + 10 y, entry q = p0*(m+1) + p1 must be x_{p0} + 10 y_{p1}.
    m = 2
    o = pyoracle.Oracle(2, 1, m, 2, True, 0)
    o.init_adaptive_intp(lambda x: x[0] + 10 * x[1], 1e10)
    c = o.get(pyoracle.FIELD_UINTP)
    pts = [o.lagr_point(p) for p in range(m + 1)]
    for p0 in range(m + 1):
        for p1 in range(m + 1):
            assert c[p0 * (m + 1) + p1] == pts[p0] + 10 * pts[p1]


def sparse_count(dim, n):
    tot = 0
    for lv in itertools.product(range(n + 1), repeat=dim):
        if sum(lv) <= n:
            tot += int(np.prod([1 if l == 0 else 2 ** (l - 1) for l in lv]))
    return tot


@pytest.mark.parametrize("dim,n,expect", [(2, 7, 576), (4, 8, 10496), (3, 5, None)])
def test_sparse_grid_counts(dim, n, expect):
    # test/test_subs.cpp (IterativeNestedLoop enumerates all level vectors) and the
    # DGSolution constructor rule: level sum <= N, 2^(l-1) supports per level >= 1
    o = pyoracle.Oracle(dim, 1, 1, n, True, n)
    assert o.num_elem == sparse_count(dim, n)
    if expect is not None:
        assert o.num_elem == expect


# --- hash key ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("dim,smax", [(1, 10), (2, 9), (3, 7), (4, 6)])
def test_hash_key_structure(dim, smax):
    # Hash::hash_key = base(S) + rank(level vector within its level-sum shell S) * 2^S
    #                + mixed-radix support index (radix 2^l per dimension, last fastest),
    # base(S) = sum_{s<S} C(d-1+s, d-1) 2^s.  Checked with exact integer binomials, and
    # the level-vector rank must be a bijection onto 0 .. C(d-1+S, d-1)-1.
    base = 0
    for s in range(smax + 1):
        nvec = math.comb(dim - 1 + s, dim - 1)
        ranks = set()
        for lv in itertools.product(range(s + 1), repeat=dim):
            if sum(lv) != s:
                continue
            ranges = [[1] if l == 0 else range(1, 2 ** l, 2) for l in lv]
            rank = None
            for sp in itertools.product(*ranges):
                key = pyoracle.hash_key(list(lv), list(sp))
                off = key - base
                assert 0 <= off < nvec * 2 ** s
                r, ind1 = divmod(off, 2 ** s)
                mr = 0
                for l, j in zip(lv, sp):
                    mr = mr * 2 ** l + (j - 1) // 2
                assert ind1 == mr
                assert rank is None or rank == r
                rank = r
            ranks.add(rank)
        assert ranks == set(range(nvec))
        base += nvec * 2 ** s


# --- 1D tables: closed forms and independent quadrature -------------------------------------------

def fine_integral(f, a, b, nsub=64):
    x, w = np.polynomial.legendre.leggauss(12)
    tot = 0.0
    edges = np.linspace(a, b, nsub + 1)
    for lo, hi in zip(edges[:-1], edges[1:]):
        t = 0.5 * (hi - lo) * x + 0.5 * (hi + lo)
        tot += 0.5 * (hi - lo) * np.sum(w * np.array([f(tt) for tt in t]))
    return tot


@pytest.mark.parametrize("k", [1, 2, 3])
def test_alpert_mass_is_identity(k):
    o = pyoracle.Oracle(1, k, 2, 4, False, 0)
    uv = o.table(4)
    assert np.max(np.abs(uv - np.eye(uv.shape[0]))) < 1e-13


def test_alpert_stiffness_independent_quadrature():
    k, nmax = 2, 3
    o = pyoracle.Oracle(1, k, 2, nmax, False, 0)
    uvx = o.table(5)
    n = uvx.shape[0]
    rng = np.random.default_rng(0)
    for r, c in rng.integers(0, n, size=(40, 2)):
        f = lambda x: o.basis_val(0, int(r), x) * o.basis_val(0, int(c), x, 1)  # noqa: E731
        assert abs(fine_integral(f, 0.0, 1.0) - uvx[r, c]) < 1e-11


def test_lagr_alpert_tables_independent_quadrature():
    o = pyoracle.Oracle(1, 2, 3, 3, False, 0)
    uv, uvx = o.table(1), o.table(2)
    rng = np.random.default_rng(1)
    for r, c in rng.integers(0, uv.shape[0], size=(30, 2)):
        c = int(c) % uv.shape[1]
        f0 = lambda x: o.basis_val(1, int(r), x) * o.basis_val(0, c, x)  # noqa: E731
        f1 = lambda x: o.basis_val(1, int(r), x) * o.basis_val(0, c, x, 1)  # noqa: E731
        assert abs(fine_integral(f0, 0.0, 1.0) - uv[r, c]) < 1e-11
        assert abs(fine_integral(f1, 0.0, 1.0) - uvx[r, c]) < 1e-10


def edge_terms_independent(o, kind_u, r, c):
    """sum over discontinuities x of v (periodic) of u(x^-)[v](x) and u(x^+)[v](x),
    with one-sided values taken 1e-9 away from x."""
    h = 1e-9
    lvl = 0 if c // 3 == 0 else int(math.log2(c // 3)) + 1
    s = c // 3
    if lvl <= 1:
        pts = [0.0, 0.5]
    else:
        pos = s - (1 << (lvl - 1))
        w = 2.0 ** -(lvl - 1)
        pts = [pos * w, pos * w + w / 2, (pos + 1) * w]
    per = lambda f, x: f(x % 1.0)  # noqa: E731
    lft = rgt = 0.0
    for x in pts:
        ul = per(lambda t: o.basis_val(kind_u, r, t), x - h if x > 0 else 1 - h)
        ur = per(lambda t: o.basis_val(kind_u, r, t), x + h if x < 1 else h)
        vl = o.basis_val(0, c, x - h if x > 0 else 1 - h)
        vr = o.basis_val(0, c, x + h if x < 1 else h)
        lft += ul * (vr - vl)
        rgt += ur * (vr - vl)
    return lft, rgt


def test_alpert_flux_tables_independent():
    o = pyoracle.Oracle(1, 2, 3, 3, False, 0)
    ul, ur, uj = o.table(6), o.table(7), o.table(8)
    assert np.array_equal(uj, ur - ul)
    rng = np.random.default_rng(2)
    for r, c in rng.integers(0, ul.shape[0], size=(60, 2)):
        lft, rgt = edge_terms_independent(o, 0, int(r), int(c))
        assert abs(lft - ul[r, c]) < 1e-5 and abs(rgt - ur[r, c]) < 1e-5


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5])
def test_lagrange_basis_interpolation_property(m):
    nmax = 4
    o = pyoracle.Oracle(1, 1, m, nmax, False, 0)
    n = (1 << nmax) * (m + 1)
    pts = [o.lagr_point(i) for i in range(n)]
    lev = [0 if i // (m + 1) == 0 else int(math.log2(i // (m + 1))) + 1 for i in range(n)]
    # degree 5 (the continuous mesh) puts points ROUND_OFF = 1e-13 inside the
    # cell ends and on both sides of the midpoint, where the nodal functions are
    # exactly 1 or 0: the values there are off by |phi'| * 1e-13 (measured up to
    # 2e-11); a wrong coefficient or factor is off by O(1)
    tol1, tol0 = (1e-13, 1e-12) if m < 5 else (1e-10, 1e-10)
    for i in range(n):
        for j in range(n):
            v = o.basis_val(1, i, pts[j])
            if j == i:
                assert abs(v - 1.0) < tol1, (i, v)
            elif lev[j] <= lev[i] and lev[j] >= 0 and (lev[j] < lev[i] or j // (m + 1) == i // (m + 1)):
                # the basis vanishes at all its own other points and at every coarser point
                assert abs(v) < tol0, (i, j, v)


@pytest.mark.parametrize("dim,m", [(2, 3), (3, 2), (2, 4), (2, 1), (2, 5)])
def test_surplus_vanishes_for_polynomials(dim, m):
    # interpolation of a polynomial of degree <= m in each variable is exact, so
    # every surplus on levels >= 1 vanishes (pins points, weights and the transform)
    o = pyoracle.Oracle(dim, 1, m, 4, True, 4)
    coef = np.random.default_rng(3).standard_normal((m + 1,) * dim)

    def f(x):
        v = 0.0
        for idx in itertools.product(range(m + 1), repeat=dim):
            v += coef[idx] * np.prod([x[d] ** idx[d] for d in range(dim)])
        return v

    o.init_adaptive_intp(f, 1e10)
    lev, _, _ = o.grid()
    c = o.get(pyoracle.FIELD_UINTP).reshape(o.num_elem, -1)
    for e in range(o.num_elem):
        if lev[e].max() > 0:
            assert np.max(np.abs(c[e])) < 1e-11


def test_alpert_projection_of_linear_function():
    # f = x on [0,1]^1 grid: c_{0,0} = 1/2, c_{0,1} = sqrt(3)/6, all finer coefficients vanish
    o = pyoracle.Oracle(1, 2, 3, 4, False, 4)
    o.init_adaptive_intp(lambda x: x[0], 1e10)
    u = o.get(pyoracle.FIELD_UCOE).reshape(o.num_elem, -1)
    lev, _, _ = o.grid()
    e0 = int(np.nonzero(lev[:, 0] == 0)[0][0])
    assert abs(u[e0, 0] - 0.5) < 1e-14 and abs(u[e0, 1] - math.sqrt(3) / 6) < 1e-14 and abs(u[e0, 2]) < 1e-14
    for e in range(o.num_elem):
        if lev[e, 0] > 0:
            assert np.max(np.abs(u[e])) < 1e-13


# --- transforms: brute force and algebraic identities ---------------------------------------------

def slots(o):
    lev, sup, _ = o.grid()
    return np.array([[0 if lev[e, d] == 0 else (1 << (lev[e, d] - 1)) + (sup[e, d] - 1) // 2
                      for d in range(o.dim)] for e in range(o.num_elem)])


@pytest.mark.parametrize("sparse", [True, False])
def test_eval_up_equals_dense_tensor_operator(sparse):
    # the L/U partial-sum algorithm is exact on downward-closed grids: it equals the
    # dense tensor-product operator sum_{e'} u(e') (x) T over ALL element pairs
    k, m, n = 2, 3, 3
    o = pyoracle.Oracle(2, k, m, n, sparse, n)
    T = o.table(0)
    rng = np.random.default_rng(4)
    u = rng.standard_normal((o.num_elem, k + 1, k + 1))
    o.set(pyoracle.FIELD_UCOE, u.ravel())
    o.eval_up()
    got = o.get(pyoracle.FIELD_UP).reshape(o.num_elem, m + 1, m + 1)
    s = slots(o)
    want = np.zeros_like(got)
    for e in range(o.num_elem):
        A = T[s[e, 0] * (k + 1):(s[e, 0] + 1) * (k + 1)]  # rows of source basis of elem e in dim 0
        B = T[s[e, 1] * (k + 1):(s[e, 1] + 1) * (k + 1)]
        for t in range(o.num_elem):
            Ta = A[:, s[t, 0] * (m + 1):(s[t, 0] + 1) * (m + 1)]
            Tb = B[:, s[t, 1] * (m + 1):(s[t, 1] + 1) * (m + 1)]
            want[t] += Ta.T @ u[e] @ Tb
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))


def test_rhs_vol_equals_dense_tensor_operator():
    k, m, n = 2, 3, 3
    o = pyoracle.Oracle(2, k, m, n, True, n)
    uv, uvx = o.table(1), o.table(2)
    rng = np.random.default_rng(5)
    f0 = rng.standard_normal((o.num_elem, m + 1, m + 1))
    f1 = rng.standard_normal((o.num_elem, m + 1, m + 1))
    o.set(10, f0.ravel())
    o.set(11, f1.ravel())
    o.rhs_zero()
    o.rhs_lagr_vol()
    got = o.get(pyoracle.FIELD_RHS).reshape(o.num_elem, k + 1, k + 1)
    s = slots(o)
    want = np.zeros_like(got)
    for t in range(o.num_elem):
        for e in range(o.num_elem):
            def blk(M, d):
                return M[s[e, d] * (m + 1):(s[e, d] + 1) * (m + 1), s[t, d] * (k + 1):(s[t, d] + 1) * (k + 1)]
            want[t] += blk(uvx, 0).T @ f0[e] @ blk(uv, 1)   # d/dx term with flux f_0
            want[t] += blk(uv, 0).T @ f1[e] @ blk(uvx, 1)   # d/dy term with flux f_1
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))


def test_linear_matrix_free_equals_assembled():
    o = pyoracle.Oracle(2, 2, 3, 5, True, 5)
    o.set(pyoracle.FIELD_UCOE, np.random.default_rng(6).standard_normal(o.num_elem * 9))
    o.rhs_zero()
    o.rhs_linear([1.0, -0.7])
    a = o.get(pyoracle.FIELD_RHS)
    o.rhs_zero()
    o.rhs_linear([1.0, -0.7], assembled=True)
    b = o.get(pyoracle.FIELD_RHS)
    assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(b))


def test_penalty_shortcut_equals_partial_sums():
    # HyperbolicAlptRHS::rhs_flx_penalty_scalar notes that the one-sweep form equals the
    # full partial-sum form with u_v (identity) in the other dimension
    k = 2
    o = pyoracle.Oracle(2, k, 3, 4, True, 4)
    u = np.random.default_rng(7).standard_normal(o.num_elem * 9)
    o.set(pyoracle.FIELD_UCOE, u)
    o.rhs_zero()
    o.rhs_penalty([1.2, 0.8])
    a = o.get(pyoracle.FIELD_RHS)
    b = np.zeros_like(a)
    sz = [k + 1, k + 1]
    for d, alpha in ((0, 1.2), (1, 0.8)):
        coef = -alpha / 2
        # ordering {0: L, 1: full} and {1: full, 0: U}, matrices ujp in d, u_v elsewhere
        for dims, lus in (((0, 1), (1, 0)), ((1, 0), (0, 2))):
            x = u
            for step, (dd, lu) in enumerate(zip(dims, lus)):
                table = 8 if dd == d else 4
                x = o.sweep(table, 1 if dd == d else 0, lu, dd, sz, sz, coef if step == 0 else 1.0, x)
            b += x
    assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(a))


def test_linear_advection_convergence():
    # u_t + u_x + u_y = 0, k=2 sparse grids: error decreases at a rate near k+1
    errs = []
    T = 0.05
    for n in (4, 5, 6):
        o = pyoracle.Oracle(2, 2, 3, n, True, n)
        o.init_adaptive_intp(lambda x: math.sin(2 * PI * (x[0] + x[1])), 1e10)
        dt = 2.0 ** -n * 0.1 / 2
        steps = math.ceil(T / dt) + 1
        dt = T / steps
        for _ in range(steps):
            o.rk_init()
            for st in range(3):
                o.rhs_zero()
                o.rhs_linear([1.0, 1.0])
                o.rk_stage(3, st, dt)
        ex = pyoracle.Oracle(2, 2, 3, n + 2, True, n + 2)
        ex.init_adaptive_intp(lambda x: math.sin(2 * PI * (x[0] + x[1] - 2 * T)), 1e-7)
        errs.append(o.l2_error_split(ex))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(o_ > 1.8 for o_ in orders), (errs, orders)


@pytest.mark.parametrize("scheme,nstage,order_expected", [(0, 1, 1), (2, 2, 2), (3, 3, 3)])
def test_rk_temporal_order(scheme, nstage, order_expected):
    # Fixed grid, dt, dt/2, dt/4: successive differences shrink by 2^order
    # (forward Euler 1, RK2 midpoint 2, SSP-RK3 3), which a wrong stage
    # coefficient breaks.  The linear operator is the one pinned above.
    n, T = 3, 0.05
    sols = []
    for nsteps in (20, 40, 80):
        o = pyoracle.Oracle(2, 2, 3, n, True, n)
        o.init_adaptive_intp(lambda x: math.sin(2 * PI * (x[0] + x[1])), 1e10)
        dt = T / nsteps
        for _ in range(nsteps):
            o.rk_init()
            for st in range(nstage):
                o.rhs_zero()
                o.rhs_linear([1.0, 1.0])
                o.rk_stage(scheme, st, dt)
        sols.append(o.get(pyoracle.FIELD_UCOE))
    r = np.linalg.norm(sols[0] - sols[1]) / np.linalg.norm(sols[1] - sols[2])
    assert abs(math.log2(r) - order_expected) < 0.15, (r, math.log2(r))
