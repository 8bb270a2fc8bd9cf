"""Shared helpers for the tests: problem definitions and reference-style loops
run on the oracle, mirroring the library's run drivers step by step."""
import ctypes

import numpy as np

import pyadg
import pyoracle

PI = np.pi


def ic_params(kind, **kw):
    p = pyadg.ICParams()
    p.kind = kind
    for name, val in kw.items():
        if name == "bumps":
            p.nbump = len(val)
            for b, (amp, width, center) in enumerate(val):
                p.bump_amp[b] = amp
                p.bump_width[b] = width
                for d, c in enumerate(center):
                    p.bump_center[b][d] = c
        elif name in ("center", "scale", "offset"):
            arr = getattr(p, name)
            for d, c in enumerate(val):
                arr[d] = c
        else:
            setattr(p, name, val)
    return p


def init_both(gpu, orc, params, eps):
    """Initialise a GPU solution and an oracle with the library's built-in
    point function so both evaluate bit-identical point values."""
    gpu.init_adaptive_intp_builtin(params, eps)
    fn = pyadg.ic_eval_ptr()
    orc._keep.append(params)
    orc.init_adaptive_intp_c(fn, ctypes.cast(ctypes.pointer(params), ctypes.c_void_p), eps)


def oracle_hyperbolic_rhs(o, prob):
    o.eval_up()
    o.flux(prob.flux_kind, (prob.flux_param,))
    o.fp_to_coe()
    o.rhs_zero()
    o.rhs_lagr_vol()
    o.rhs_lagr_flx()
    o.rhs_penalty([prob.alpha[d] for d in range(o.dim)])


def oracle_run_hyperbolic(o, prob):
    """The loop of adg_run_hyperbolic (reference example/02_hyperbolic_05_burgers_adapt.cpp:223-498)."""
    t, steps, dof_stages = 0.0, 0, 0
    while t < prob.t_final and (prob.max_steps <= 0 or steps < prob.max_steps):
        lm = o.max_level()
        s = 0.0
        for d in range(o.dim):
            s += abs(prob.wave_speed[d]) * 2.0 ** int(lm[d])
        dt = prob.cfl / s
        dt = min(dt, prob.t_final - t)
        if prob.adaptive:
            o.copy_to_predict()
            o.rk_init()
            oracle_hyperbolic_rhs(o, prob)
            o.rk_stage(0, 0, dt)
            o.refine(prob.eps)
            o.copy_predict_to_ucoe()
        o.rk_init()
        for st in range(3):
            oracle_hyperbolic_rhs(o, prob)
            o.rk_stage(3, st, dt)
        dof_stages += 3 * o.num_elem * o.K
        if prob.adaptive:
            o.coarsen(prob.eta)
        t += dt
        steps += 1
    return steps, dof_stages, t


def oracle_run_linear(o, coef, dt, nsteps):
    for _ in range(nsteps):
        o.rk_init()
        for st in range(3):
            o.rhs_zero()
            o.rhs_linear(coef)
            o.rk_stage(3, st, dt)


def problem(**kw):
    p = pyadg.Problem()
    for name, val in kw.items():
        if name in ("alpha", "wave_speed"):
            arr = getattr(p, name)
            for d, c in enumerate(val):
                arr[d] = c
        else:
            setattr(p, name, val)
    return p


def same_grid(a, b):
    la, sa, ka = a.grid()
    lb, sb, kb = b.grid()
    return np.array_equal(la, lb) and np.array_equal(sa, sb) and np.array_equal(ka, kb)


def first_diff(a, b):
    idx = np.nonzero(a != b)[0]
    if idx.size == 0:
        return "identical"
    i = idx[0]
    return f"{idx.size} differ; first at {i}: {a[i]!r} vs {b[i]!r}; max abs {np.max(np.abs(a - b)):.3e}"
