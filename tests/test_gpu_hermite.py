"""Hermite P3 flux interpolation on the GPU against the CPU oracle
(HermInterpolation::nonlinear_Herm_2D_fast with FastHermIntp and
HyperbolicSameFluxHermRHS / HyperbolicDiffFluxHermRHS; the reference's
Burgers and KPP drivers, example/02_hyperbolic_05_burgers_adapt.cpp:192-276,
example/02_hyperbolic_07_kpp.cpp:201-276).

Burgers runs are bit-exact; sin/cos fluxes differ from the host libm by a few
ulps, so those compare with a relative tolerance of 1e-12."""
import numpy as np
import pytest

import pyadg
import pyoracle
from common import first_diff, ic_params, init_both, oracle_hyperbolic_rhs, oracle_run_hyperbolic, problem, same_grid

pytestmark = pytest.mark.gpu


def hermite_pair(dim, k, nmax, level_init, params, eps):
    print(f"[case] hermite dim={dim} k={k} nmax={nmax} level_init={level_init}", flush=True)
    g = pyadg.Solution(dim, k, 3, nmax, True, level_init)
    o = pyoracle.Oracle(dim, k, 3, nmax, True, level_init)
    # the initialisation runs on the Lagrange family whatever the family in use
    g.set_interpolation(pyadg.INTERP_HERMITE)
    o.set_interpolation(1)
    init_both(g, o, params, eps)
    assert same_grid(g, o), "adaptive interpolation produced different element sets"
    return g, o


def burgers_ic(dim):
    return ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6)[:dim]), (0.1, 0.1, (0.7, 0.3)[:dim])])


def sincos_ic():
    # smooth data: the KPP driver's discontinuous data need the artificial viscosity the reference adds
    # there (slopes of a jump resolved at the finest level blow the Hermite flux up within two steps)
    return ic_params(2, base=0.5, ampl=0.5, bumps=[(1.5, 0.1, (0.5, 0.5))])


def test_hermite_operators_bitexact_burgers():
    g, o = hermite_pair(2, 2, 6, 4, burgers_ic(2), 1e-4)
    g.eval_up(); o.eval_up()
    a, b = g.get(pyadg.FIELD_UP), o.get(pyoracle.FIELD_UP)
    assert np.array_equal(a, b), "eval_up: " + first_diff(a, b)
    g.flux(pyadg.FLUX_BURGERS); o.flux(0)
    g.fp_to_coe(); o.fp_to_coe()
    for d in range(2):
        a, b = g.get(pyadg.FIELD_FUCOE + d), o.get(pyoracle.FIELD_FUCOE + d)
        assert np.array_equal(a, b), f"fucoe[{d}]: " + first_diff(a, b)
    g.rhs_zero(); o.rhs_zero()
    g.rhs_lagr_vol(); o.rhs_lagr_vol()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "rhs vol: " + first_diff(a, b)
    g.rhs_lagr_flx(); o.rhs_lagr_flx()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "rhs flx: " + first_diff(a, b)
    g.rhs_penalty([1.2, 1.2]); o.rhs_penalty([1.2, 1.2])
    g.rk_init(); o.rk_init()
    for st in range(3):
        g.rk_stage(pyadg.RK3, st, 1e-3); o.rk_stage(3, st, 1e-3)
        a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, b), f"rk stage {st}: " + first_diff(a, b)


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("form", [0, 1])
def test_hermite_adaptive_burgers_run_bitexact(dim, form):
    g, o = hermite_pair(dim, 2, 6, 4, burgers_ic(dim), 1e-4)
    prob = problem(flux_kind=0, flux_param=0.0, alpha=[1.2] * dim, wave_speed=[1.0] * dim, cfl=0.2, t_final=0.03,
                   adaptive=1, eps=1e-4, eta=1e-5, max_steps=5, form=form)
    st = g.run_hyperbolic(prob)
    steps, dof_stages, _ = oracle_run_hyperbolic(o, prob)
    assert st.steps == steps and st.dof_stages == dof_stages
    assert same_grid(g, o), "element sets differ after adaptive steps"
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.array_equal(a, b), first_diff(a, b)


@pytest.mark.parametrize("form", [0, 1])
def test_hermite_sincos_flux_tolerance(form):
    g, o = hermite_pair(2, 2, 6, 4, sincos_ic(), 1e-3)
    prob = problem(flux_kind=1, flux_param=1.0, alpha=[1.0, 1.0], wave_speed=[1.0, 1.0], cfl=0.1, t_final=1.0,
                   adaptive=0, eps=1e-3, eta=1e-4, max_steps=3, form=form)
    g.run_hyperbolic(prob)
    oracle_run_hyperbolic(o, prob)
    a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, -1])
def test_hermite_fused_stage_matches_per_sweep_path(kind, variant):
    p = burgers_ic(2) if kind == 0 else sincos_ic()
    prob = problem(flux_kind=kind, flux_param=1.0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=1.0)
    a, _ = hermite_pair(2, 2, 6, 4, p, 1e-4)
    b = pyadg.Solution(2, 2, 3, 6, True, 4)
    b.init_adaptive_intp_builtin(p, 1e-4)
    b.set_interpolation(pyadg.INTERP_HERMITE)
    b.set_kernel_variant(variant)
    assert same_grid(a, b)
    a.rk_init(); b.rk_init()
    for st in range(3):
        a.eval_up(); a.flux(kind, (1.0,)); a.fp_to_coe(); a.rhs_zero(); a.rhs_lagr_vol()
        a.rhs_lagr_flx(); a.rhs_penalty([1.2, 1.2]); a.rk_stage(pyadg.RK3, st, 1e-3)
        b.stage(prob, pyadg.RK3, st, 1e-3)
        for f in (pyadg.FIELD_UP, pyadg.FIELD_RHS, pyadg.FIELD_UCOE):
            x, y = a.get(f), b.get(f)
            assert np.array_equal(x, y), f"stage {st} field {f}: " + first_diff(x, y)


def test_switching_families_keeps_each_bitexact():
    # Hermite stage, then Lagrange stage on the same context, against the oracle doing the same
    g, o = hermite_pair(2, 2, 6, 4, burgers_ic(2), 1e-4)
    prob = problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=1.0)
    g.rk_init(); o.rk_init()
    for st, fam in enumerate((1, 0, 1)):
        g.set_interpolation(fam); o.set_interpolation(fam)
        g.stage(prob, pyadg.RK3, st, 1e-3)
        oracle_hyperbolic_rhs(o, prob); o.rk_stage(3, st, 1e-3)
        a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, b), f"stage {st} family {fam}: " + first_diff(a, b)


def test_hermite_rejects_what_the_reference_cannot_do():
    print("[case] hermite argument checks", flush=True)
    s = pyadg.Solution(2, 2, 2, 4, True, 4)
    with pytest.raises(RuntimeError, match="m = 3"):
        s.set_interpolation(pyadg.INTERP_HERMITE)
    s = pyadg.Solution(3, 1, 3, 3, True, 3)
    s.set_interpolation(pyadg.INTERP_HERMITE)
    s.eval_up()
    with pytest.raises(RuntimeError, match="one- or two-dimensional"):
        s.flux(pyadg.FLUX_BURGERS)
    s = pyadg.Solution(2, 1, 3, 3, True, 3)
    s.set_interpolation(pyadg.INTERP_HERMITE)
    with pytest.raises(RuntimeError, match="flux of u alone"):
        s.flux(pyadg.FLUX_VARCOEF, (0.5,))
