"""GPU parity: the HIP library against the CPU oracle on the same inputs.

FP64 results are compared bit for bit (both sum every output in ascending
source-slot order with no FMA).  The one exception is the sin/cos flux, where
the device libm and glibc may differ in the last bit: there the tolerance is
1e-12 relative to the field's max norm.
"""
import ctypes

import numpy as np
import pytest

import pyadg
import pyoracle
from common import (first_diff, ic_params, init_both, oracle_hyperbolic_rhs, oracle_run_hyperbolic, oracle_run_linear,
                    problem, same_grid)

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(1234)


def pair(dim, k, m, nmax, sparse=True, level_init=None):
    print(f"[case] dim={dim} k={k} m={m} nmax={nmax} sparse={sparse} level_init={level_init}", flush=True)
    g = pyadg.Solution(dim, k, m, nmax, sparse, level_init)
    o = pyoracle.Oracle(dim, k, m, nmax, sparse, level_init)
    assert same_grid(g, o)
    return g, o


def adapted_pair(dim, k, m, nmax, level_init, params, eps):
    g, o = pair(dim, k, m, nmax, True, level_init)
    init_both(g, o, params, eps)
    assert same_grid(g, o), "adaptive interpolation produced different element sets"
    return g, o


SWEEP_CASES = []
for table in range(9):
    for lu in (0, 1, 2):
        for flx in (0, 1):
            if table in (1, 2) and flx:
                continue
            SWEEP_CASES.append((table, lu, flx))


@pytest.mark.parametrize("dim,k,m,nmax", [(2, 2, 3, 5), (3, 1, 2, 4), (2, 3, 4, 4)])
def test_sweep_bitexact_sparse(dim, k, m, nmax):
    g, o = pair(dim, k, m, nmax)
    E = g.num_elem
    for table, lu, flx in SWEEP_CASES:
        pf = k if (table == 0 or table >= 4) else m
        pt = m if table == 0 else k
        for d in range(dim):
            si = [pf + 1] * dim
            so = list(si)
            so[d] = pt + 1
            x = rng.standard_normal(E * int(np.prod(si)))
            a = g.sweep(table, flx, lu, d, si, so, 0.7, x)
            b = o.sweep(table, flx, lu, d, si, so, 0.7, x)
            assert np.array_equal(a, b), f"table={table} lu={lu} flx={flx} d={d}: {first_diff(a, b)}"


def test_sweep_bitexact_adaptive_grid():
    # an adaptively refined 2D grid: fibers with holes, uneven depth
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.2, 0.06, (0.3, 0.7)), (0.1, 0.08, (0.8, 0.2))])
    g, o = adapted_pair(2, 2, 3, 7, 3, p, 1e-4)
    E = g.num_elem
    assert E > 200
    for table, lu, flx in SWEEP_CASES:
        pf = 2 if (table == 0 or table >= 4) else 3
        pt = 3 if table == 0 else 2
        for d in range(2):
            si = [pf + 1] * 2
            so = list(si)
            so[d] = pt + 1
            x = rng.standard_normal(E * int(np.prod(si)))
            a = g.sweep(table, flx, lu, d, si, so, -1.3, x)
            b = o.sweep(table, flx, lu, d, si, so, -1.3, x)
            assert np.array_equal(a, b), f"table={table} lu={lu} flx={flx} d={d}: {first_diff(a, b)}"


@pytest.mark.parametrize("m", [3, 5])
def test_init_and_operators_bitexact_burgers(m):
    # m = 5: Lagrange P5 on the continuous mesh (points ROUND_OFF inside the
    # cell ends and the midpoint), as the reference's linear-advection driver uses
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.07, (0.4, 0.6))])
    g, o = adapted_pair(2, 2, m, 6, 4, p, 1e-4)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), "init ucoe: " + first_diff(a, b)
    g.eval_up(); o.eval_up()
    a, b = g.get(pyadg.FIELD_UP), o.get(pyoracle.FIELD_UP)
    assert np.array_equal(a, b), "eval_up: " + first_diff(a, b)
    g.flux(pyadg.FLUX_BURGERS); o.flux(0)
    g.fp_to_coe(); o.fp_to_coe()
    for d in range(2):
        a, b = g.get(pyadg.FIELD_FUCOE + d), o.get(10 + d)
        assert np.array_equal(a, b), f"fucoe[{d}]: " + first_diff(a, b)
    g.rhs_zero(); o.rhs_zero()
    g.rhs_lagr_vol(); o.rhs_lagr_vol()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "rhs vol: " + first_diff(a, b)
    g.rhs_lagr_flx(); o.rhs_lagr_flx()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "rhs flx: " + first_diff(a, b)
    g.rhs_penalty([1.2, 1.2]); o.rhs_penalty([1.2, 1.2])
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "rhs penalty: " + first_diff(a, b)
    g.rk_init(); o.rk_init()
    for st in range(3):
        g.rk_stage(pyadg.RK3, st, 1e-3); o.rk_stage(3, st, 1e-3)
        a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, b), f"rk stage {st}: " + first_diff(a, b)


def test_linear_run_bitexact():
    # config-1 shape at a reduced level: 2D, k=2, sparse N=6, RK3 steps of u_t + u_x + u_y = 0
    p = ic_params(0, shift=0.0)
    g, o = pair(2, 2, 3, 6)
    init_both(g, o, p, 1e10)
    dt = 2.0 ** -6 * 0.1 / 2
    st = g.run_linear([1.0, 1.0], dt, 12)
    oracle_run_linear(o, [1.0, 1.0], dt, 12)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)
    assert st.dof_stages == 12 * 3 * g.num_elem * 9


def test_refine_coarsen_bitexact():
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.2, 0.05, (0.25, 0.75))])
    g, o = adapted_pair(2, 2, 3, 7, 3, p, 1e-3)
    x = g.get(pyadg.FIELD_UCOE)
    # perturb the coefficients so that both tests fire on many elements
    y = x * (1.0 + 0.5 * rng.standard_normal(x.size))
    g.set(pyadg.FIELD_UCOE, y); o.set(pyoracle.FIELD_UCOE, y)
    na = g.refine(1e-4); nb = o.refine(1e-4)
    assert na == nb and same_grid(g, o)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)
    ra = g.coarsen(3e-4); rb = o.coarsen(3e-4)
    assert ra == rb and same_grid(g, o)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


def test_adaptive_burgers_run_bitexact():
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6)), (0.1, 0.1, (0.7, 0.3))])
    g, o = adapted_pair(2, 2, 3, 6, 4, p, 1e-4)
    prob = problem(flux_kind=0, flux_param=0.0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=0.02,
                   adaptive=1, eps=1e-4, eta=1e-5, max_steps=4)
    st = g.run_hyperbolic(prob)
    steps, dof_stages, t = oracle_run_hyperbolic(o, prob)
    assert st.steps == steps and st.dof_stages == dof_stages
    assert same_grid(g, o), "element sets differ after adaptive steps"
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


def test_varcoef_4d_bitexact():
    p = ic_params(1, shift=0.0)
    g, o = pair(4, 1, 2, 4)
    init_both(g, o, p, 1e10)
    prob = problem(flux_kind=2, flux_param=0.5, alpha=[1.5] * 4, wave_speed=[1.5] * 4, cfl=0.2, t_final=1.0,
                   adaptive=0, eps=0.0, eta=0.0, max_steps=2)
    st = g.run_hyperbolic(prob)
    oracle_run_hyperbolic(o, prob)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)
    assert st.steps == 2


def test_sincos_flux_tolerance():
    p = ic_params(3, inside=7 * np.pi / 4, outside=np.pi / 4, radius=1.0, center=[0.0, 0.0],
                  scale=[4.0, 4.0], offset=[-2.0, -2.5])
    g, o = adapted_pair(2, 2, 3, 6, 4, p, 1e-3)
    prob = problem(flux_kind=1, flux_param=0.25, alpha=[0.25, 0.25], wave_speed=[0.25, 0.25], cfl=0.2, t_final=1.0,
                   adaptive=0, eps=1e-3, eta=1e-4, max_steps=3)
    g.run_hyperbolic(prob)
    oracle_run_hyperbolic(o, prob)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("dim,k,m,nmax,level_init,kind", [(2, 2, 3, 6, 4, 0), (2, 2, 3, 6, 4, 1), (3, 1, 2, 4, 3, 2),
                                                          (4, 1, 2, 4, 4, 2), (2, 3, 4, 5, 4, 2)])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, -1])
def test_fused_stage_matches_per_sweep_path(dim, k, m, nmax, level_init, kind, variant):
    # the fused stage must reproduce the per-sweep kernels bit for bit (same sums, same order)
    if kind == 0:
        p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6))])
        prob = problem(flux_kind=0, alpha=[1.2] * dim, wave_speed=[1.0] * dim, cfl=0.2, t_final=1.0)
    elif kind == 1:
        p = ic_params(3, inside=7 * np.pi / 4, outside=np.pi / 4, radius=1.0, center=[0.0, 0.0], scale=[4.0, 4.0],
                      offset=[-2.0, -2.5])
        prob = problem(flux_kind=1, flux_param=0.25, alpha=[0.25] * dim, wave_speed=[0.25] * dim, cfl=0.2, t_final=1.0)
    else:
        p = ic_params(1, shift=0.1)
        prob = problem(flux_kind=2, flux_param=0.5, alpha=[1.5] * dim, wave_speed=[1.5] * dim, cfl=0.2, t_final=1.0)
    a, _ = adapted_pair(dim, k, m, nmax, level_init, p, 1e-4)
    b = pyadg.Solution(dim, k, m, nmax, True, level_init)
    b.init_adaptive_intp_builtin(p, 1e-4)
    b.set_kernel_variant(variant)
    assert same_grid(a, b)
    a.rk_init(); b.rk_init()
    for st in range(3):
        a.eval_up(); a.flux(prob.flux_kind, (prob.flux_param,)); a.fp_to_coe(); a.rhs_zero(); a.rhs_lagr_vol()
        a.rhs_lagr_flx(); a.rhs_penalty([prob.alpha[d] for d in range(dim)]); a.rk_stage(pyadg.RK3, st, 1e-3)
        b.stage(prob, pyadg.RK3, st, 1e-3)
        for f in (pyadg.FIELD_UP, pyadg.FIELD_RHS, pyadg.FIELD_UCOE):
            x, y = a.get(f), b.get(f)
            assert np.array_equal(x, y), f"stage {st} field {f}: " + first_diff(x, y)


def test_fused_adaptive_burgers_run_bitexact():
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6)), (0.1, 0.1, (0.7, 0.3))])
    g, o = adapted_pair(2, 2, 3, 6, 4, p, 1e-4)
    prob = problem(flux_kind=0, flux_param=0.0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=0.03,
                   adaptive=1, eps=1e-4, eta=1e-5, max_steps=6, form=1)
    st = g.run_hyperbolic(prob)
    steps, dof_stages, t = oracle_run_hyperbolic(o, prob)
    assert st.steps == steps and st.dof_stages == dof_stages
    assert same_grid(g, o), "element sets differ after adaptive steps"
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


def test_rk2_midpoint_stages_bitexact():
    # RK2Midpoint::step_stage (source/ODESolver.cpp:247-259) on the per-sweep and the fused path
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6))])
    g, o = adapted_pair(2, 2, 3, 6, 4, p, 1e-4)
    f = pyadg.Solution(2, 2, 3, 6, True, 4)
    f.init_adaptive_intp_builtin(p, 1e-4)
    prob = problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=1.0)
    g.rk_init(); o.rk_init(); f.rk_init()
    for st in range(2):
        g.eval_up(); g.flux(0); g.fp_to_coe(); g.rhs_zero(); g.rhs_lagr_vol(); g.rhs_lagr_flx()
        g.rhs_penalty([1.2, 1.2]); g.rk_stage(pyadg.RK2, st, 2e-3)
        oracle_hyperbolic_rhs(o, prob); o.rk_stage(2, st, 2e-3)
        f.stage(prob, pyadg.RK2, st, 2e-3)
        a, b, c = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE), f.get(pyadg.FIELD_UCOE)
        assert np.array_equal(a, b), f"rk2 stage {st}: " + first_diff(a, b)
        assert np.array_equal(c, b), f"fused rk2 stage {st}: " + first_diff(c, b)


@pytest.mark.parametrize("dim,k,m,nmax,cap", [(2, 2, 3, 6, 4), (3, 1, 2, 4, 3)])
def test_coarse_grid_cap_bitexact(dim, k, m, nmax, cap):
    # the _coarse_grid variants: per-sweep path and fused stage against the oracle
    p = ic_params(1, shift=0.1)
    g, o = pair(dim, k, m, nmax)
    init_both(g, o, p, 1e10)
    f = pyadg.Solution(dim, k, m, nmax, True, nmax)
    f.init_adaptive_intp_builtin(p, 1e10)
    for s in (g, o, f):
        s.set_level_sum_cap(cap)
    g.eval_up(); o.eval_up()
    a, b = g.get(pyadg.FIELD_UP), o.get(pyoracle.FIELD_UP)
    assert np.array_equal(a, b), "eval_up: " + first_diff(a, b)
    prob = problem(flux_kind=2, flux_param=0.5, alpha=[1.5] * dim, wave_speed=[1.5] * dim, cfl=0.2, t_final=1.0)
    g.rk_init(); o.rk_init(); f.rk_init()
    for st in range(2):
        g.eval_up(); g.flux(2, (0.5,)); g.fp_to_coe(); g.rhs_zero(); g.rhs_lagr_vol(); g.rhs_lagr_flx()
        g.rhs_penalty([1.5] * dim); g.rk_stage(pyadg.RK2, st, 1e-3)
        oracle_hyperbolic_rhs(o, prob); o.rk_stage(2, st, 1e-3)
        f.stage(prob, pyadg.RK2, st, 1e-3)
        a, b, c = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE), f.get(pyadg.FIELD_UCOE)
        assert np.array_equal(a, b), f"stage {st}: " + first_diff(a, b)
        assert np.array_equal(c, b), f"fused stage {st}: " + first_diff(c, b)
    coef = [1.0, -0.5, 0.75][:dim]
    g.rhs_zero(); o.rhs_zero()
    g.rhs_linear(coef); o.rhs_linear(coef)
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "linear: " + first_diff(a, b)


@pytest.mark.parametrize("dim,k,nmax,level_init,kind,nfunc,eps", [(2, 2, 7, 2, 0, 2, 1e-4), (4, 1, 6, 6, 1, 1, 1e10),
                                                                   (3, 2, 6, 1, 1, 1, 1e-3)])
def test_separable_init_bitexact(dim, k, nmax, level_init, kind, nfunc, eps):
    # DGAdapt::init_separable_scalar_sum with refinement, against the oracle
    p = ic_params(kind, shift=0.05)
    g, o = pair(dim, k, k + 1, nmax, True, level_init)
    g.init_separable_builtin(p, nfunc, eps)
    o._keep.append(p)
    o.init_separable_c(pyadg.sep_eval_ptr(), ctypes.cast(ctypes.pointer(p), ctypes.c_void_p), nfunc, eps)
    assert same_grid(g, o), "element sets differ"
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


@pytest.mark.parametrize("dim", [2, 3])
def test_filters_bitexact(dim):
    p = ic_params(1, shift=0.1)
    g, o = pair(dim, 1, 2, 5)
    init_both(g, o, p, 1e10)
    g.filter(0.9, [1.0, 0.5, 2.0][:dim], 2e-3, 3)
    o.filter(0.9, [1.0, 0.5, 2.0][:dim], 2e-3, 3)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), "filter: " + first_diff(a, b)

    def speed(x, d):
        return -x[(d + 1) % dim] + 0.5

    g.filter_local(0.9, speed, 2e-3, 2)
    o.filter_local(0.9, speed, 2e-3, 2)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), "filter_local: " + first_diff(a, b)


def test_point_values_and_quadrature_error():
    # DGSolution::val on a tensor grid (boundaries included) and get_error_no_separable_scalar;
    # element contributions are summed in another order than the oracle's map order: 1e-13 relative
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6))])
    g, o = adapted_pair(2, 2, 3, 6, 3, p, 1e-4)
    x1d = np.concatenate([np.linspace(0.0, 1.0, 65), np.array([0.3141, 0.77])])
    a, b = g.eval_tensor(x1d), o.eval_tensor(x1d)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b)), np.max(np.abs(a - b))
    q = ic_params(0, shift=0.0)
    for pts in (1, 3):
        ea = g.error_no_separable_builtin(q, pts)
        eb = o.error_no_separable_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(q), ctypes.c_void_p), pts)
        assert np.all(np.abs(ea - eb) <= 1e-12 * np.abs(eb)), (ea, eb)


def test_fused_stage_refuses_lagrange_degree_5():
    # the fused kernels are instantiated for Lagrange degrees 1..4; degree 5
    # reports an error instead of launching (the per-sweep path serves it)
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.07, (0.4, 0.6))])
    g = pyadg.Solution(2, 2, 5, 5, True, 3)
    g.init_adaptive_intp_builtin(p, 1e-4)
    prob = problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=1.0)
    g.rk_init()
    with pytest.raises(pyadg.AdgError, match="unsupported basis sizes"):
        g.stage(prob, pyadg.RK3, 0, 1e-3)
