"""Every benchmark configuration on the GPU at its own size, against the CPU
oracle (reference data structures) on the same inputs.

FP64 results are bit-identical (same sums in the same order, no FMA), except
the sin/cos flux of configuration 5, where device and host libm may differ in
the last bit (coefficients within 1e-12 of the field's max norm).
"""
import ctypes
import math

import numpy as np
import pytest

import pyadg
import pyoracle
import workloads
from common import first_diff, oracle_run_hyperbolic, oracle_run_linear

pytestmark = pytest.mark.gpu


def oracle_for(w):
    o = pyoracle.Oracle(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    o._keep.append(w.ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(w.ic), ctypes.c_void_p), w.init_eps)
    return o


def same_set(g, o):
    lg, sg, kg = g.grid()
    lo, so, ko = o.grid()
    return np.array_equal(kg, ko) and np.array_equal(lg, lo) and np.array_equal(sg, so)


def test_c1_full_run_bitexact_and_error():
    w = workloads.c1_linear()
    print(f"[case] c1: sparse N=7, {w.nsteps} RK3 steps to t=1", flush=True)
    g = workloads.make_solution(w)
    o = oracle_for(w)
    assert g.num_elem == 576 and same_set(g, o)
    st = g.run_linear(w.coef, w.dt, w.nsteps)
    oracle_run_linear(o, w.coef, w.dt, w.nsteps)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)
    assert st.steps == w.nsteps == 2561
    t = w.dt * w.nsteps
    exact = pyadg.ICParams()
    exact.kind, exact.shift = 0, 2.0 * t  # u = sin(2 pi (x + y - 2 t))
    err = g.error_no_separable_builtin(exact, 3)
    print(f"[c1] L1, L2, Linf error at t={t}: {err}", flush=True)
    assert err[1] < 1e-4


# refine / coarsen on the host (0) or on the device (1, adg_set_device_adapt)
ADAPT = [pytest.param(0, id="host-adapt"), pytest.param(1, id="device-adapt")]


@pytest.mark.parametrize("adapt", ADAPT)
@pytest.mark.parametrize("seed", [1, 2])
def test_c2_full_solve_bitexact(seed, adapt):
    w = workloads.c2_burgers(seed=seed)
    print(f"[case] c2: seed {seed}, NMAX 8, N0 6, t=0.1, adapt {adapt}", flush=True)
    g = workloads.make_solution(w)
    g.set_device_adapt(adapt)
    o = oracle_for(w)
    assert same_set(g, o)
    w.problem.form = 1
    st = g.run_hyperbolic(w.problem)
    steps, dof_stages, t = oracle_run_hyperbolic(o, w.problem)
    assert st.steps == steps and st.dof_stages == dof_stages and abs(st.final_time - 0.1) < 1e-15
    assert same_set(g, o), "element sets differ"
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


def test_c3_full_grid_steps_bitexact():
    w = workloads.c3_varcoef_4d()
    print("[case] c3: 4D sparse N=8, 3 steps", flush=True)
    g = workloads.make_solution(w)
    o = oracle_for(w)
    assert g.num_elem == 10496 and same_set(g, o)
    w.problem.form, w.problem.max_steps = 1, 3
    st = g.run_hyperbolic(w.problem)
    steps, _, _ = oracle_run_hyperbolic(o, w.problem)
    assert st.steps == steps == 3
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


@pytest.mark.parametrize("adapt", ADAPT)
def test_c4_full_size_instances_bitexact(adapt):
    ws = [workloads.c4_advection_3d(seed=s) for s in (1, 3, 5, 8)]
    print(f"[case] c4: 4 instances at NMAX 7, 5 adaptive steps, adapt {adapt}", flush=True)
    b = workloads.make_batch(ws)
    b.set_device_adapt(adapt)
    prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
    prob.form, prob.max_steps = 1, 5
    b.run_hyperbolic(prob)
    assert b.device_adapt_stats() == ((10, 0) if adapt else (0, 0))
    for i, w in enumerate(ws):
        o = oracle_for(w)
        p = w.oracle_problem()
        p.max_steps = 5
        oracle_run_hyperbolic(o, p)
        lv, sp, hk = o.grid()
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv), f"instance {i}: element sets differ"
        a, c = b.instance_get(i), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, c), f"instance {i}: " + first_diff(a, c)


@pytest.mark.parametrize("adapt", ADAPT)
def test_c5_full_size_adaptive_steps(adapt):
    w = workloads.c5_rotating(seed=1)
    print(f"[case] c5: NMAX 9, 5 adaptive steps, adapt {adapt}", flush=True)
    g = workloads.make_solution(w)
    g.set_device_adapt(adapt)
    o = oracle_for(w)
    assert same_set(g, o)
    w.problem.form, w.problem.max_steps = 1, 5
    st = g.run_hyperbolic(w.problem)
    steps, _, _ = oracle_run_hyperbolic(o, w.problem)
    assert st.steps == steps == 5
    assert same_set(g, o), "element sets differ"
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def _bell(x):
    # cosine bell of the solid-body rotation test (Guo & Cheng, example 4.2), as in
    # example/02_hyperbolic_02_var_coeff.cpp (commented initial function)
    r = math.hypot(x[0] - 0.75, x[1] - 0.5)
    return 0.23 * math.cos(math.pi * r / (2 * 0.23)) ** 6 if r <= 0.23 else 0.0


def _bell_exact(x, y, t):
    # u_t + (0.5 - y) u_x + (x - 0.5) u_y = 0 rotates about (0.5, 0.5) with unit angular speed
    c, s = math.cos(t), math.sin(t)
    xs, ys = x - 0.5, y - 0.5
    x0, y0 = c * xs + s * ys + 0.5, -s * xs + c * ys + 0.5
    r = np.hypot(x0 - 0.75, y0 - 0.5)
    return np.where(r <= 0.23, 0.23 * np.cos(np.pi * np.minimum(r, 0.23) / 0.46) ** 6, 0.0)


def _solid_body_run(n, tf, oracle=False):
    """Sparse grid N, k=2, Lagrange P3 interpolation of a(x) u computed on the
    host (the reference driver's coe_func closure), LF penalty alpha 0.5,
    SSP-RK3, dt = 0.2 / (2^N + 2^N)."""
    k, m = 2, 3
    g = pyadg.Solution(2, k, m, n, True, n)
    g.init_adaptive_intp(_bell, 1e10)
    o = None
    if oracle:
        # same sparse grid (no refinement at eps 1e10), same initial coefficients
        o = pyoracle.Oracle(2, k, m, n, True, n)
        o.set(pyoracle.FIELD_UCOE, g.get(pyadg.FIELD_UCOE))
    x = g.lagr_points()
    a0 = (0.5 - x[:, :, 1]).ravel()
    a1 = (x[:, :, 0] - 0.5).ravel()
    dt_full = 0.2 / (2.0 ** n + 2.0 ** n)
    t = 0.0
    while t < tf:
        dt = min(dt_full, tf - t)
        for s in ([g] + ([o] if o is not None else [])):
            s.rk_init()
        for st in range(3):
            for s in ([g] + ([o] if o is not None else [])):
                s.eval_up()
                up = s.get(pyadg.FIELD_UP if s is g else pyoracle.FIELD_UP)
                fp = np.concatenate([a0 * up, a1 * up])
                if s is g:
                    s.flux_set(fp, 2)
                else:
                    s.set(20, fp[:len(up)])
                    s.set(21, fp[len(up):])
                s.fp_to_coe()
                s.rhs_zero()
                s.rhs_lagr_vol()
                s.rhs_lagr_flx()
                s.rhs_penalty([0.5, 0.5])
                s.rk_stage(3, st, dt)
        t += dt
    return g, o, t


def test_solid_body_rotation_convergence_orders():
    # the interpolated-flux path (FastLagrIntp + HyperbolicLagrRHS + RK3) on the
    # problem of the reference's tabulated sparse-grid runs
    # (example/02_hyperbolic_02_var_coeff.cpp:55-69: L2 orders 2.08, 2.85, 2.40,
    # 2.67, 2.36 for N = 5..10)
    tf = 0.5
    errs = []
    for n in (5, 6, 7, 8):
        print(f"[case] solid-body rotation sparse N={n}", flush=True)
        g, _, t = _solid_body_run(n, tf)
        ng = 3
        nc = 2 ** n
        xq, wq = np.polynomial.legendre.leggauss(ng)
        x1 = ((np.arange(nc)[:, None] + 0.5 * (xq[None, :] + 1)) / nc).ravel()
        w1 = np.tile(wq / (2 * nc), nc)
        uh = np.zeros(len(x1) ** 2)
        g._check(pyadg.lib().adg_eval_tensor(g._h, len(x1), pyadg._dp(x1), pyadg._dp(uh)))
        X, Y = np.meshgrid(x1, x1, indexing="ij")
        ue = _bell_exact(X, Y, t)
        W = np.outer(w1, w1)
        errs.append(math.sqrt(np.sum(W * (uh.reshape(X.shape) - ue) ** 2)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    print(f"[solid-body] L2 errors {errs}, orders {orders}", flush=True)
    assert all(1.8 <= q <= 3.4 for q in orders), orders
    assert np.mean(orders) >= 2.0


def test_solid_body_rotation_matches_oracle():
    g, o, _ = _solid_body_run(5, 0.05, oracle=True)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)
