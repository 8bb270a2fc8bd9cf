"""Reference-style time loops driven on the oracle (test infrastructure only).

They mirror the library's run drivers step by step: the loop of
example/02_hyperbolic_05_burgers_adapt.cpp:223-498 (predict, refine, RK3,
coarsen) and of example/02_hyperbolic_01_scalar_const_coefficient.cpp:164-179.
"""


def hyperbolic_rhs(o, prob):
    o.eval_up()
    # flux parameters: (s,) or, for the scaled variable coefficient, (s, c_0, .., c_{dim-1})
    o.flux(prob.flux_kind, getattr(prob, "flux_params", (prob.flux_param,)))
    o.fp_to_coe()
    o.rhs_zero()
    o.rhs_lagr_vol()
    o.rhs_lagr_flx()
    o.rhs_penalty([prob.alpha[d] for d in range(o.dim)])


def run_hyperbolic(o, prob):
    t, steps, dof_stages = prob.t_start, 0, 0
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
            hyperbolic_rhs(o, prob)
            o.rk_stage(0, 0, dt)
            o.refine(prob.eps)
            o.copy_predict_to_ucoe()
        o.rk_init()
        for st in range(3):
            hyperbolic_rhs(o, prob)
            o.rk_stage(3, st, dt)
        dof_stages += 3 * o.num_elem * o.K
        if prob.adaptive:
            o.coarsen(prob.eta)
        t += dt
        steps += 1
    return steps, dof_stages, t


def run_linear(o, coef, dt, nsteps):
    for _ in range(nsteps):
        o.rk_init()
        for st in range(3):
            o.rhs_zero()
            o.rhs_linear(coef)
            o.rk_stage(3, st, dt)
    return nsteps * 3 * o.num_elem * o.K
