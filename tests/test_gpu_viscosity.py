"""Artificial viscosity in its matrix-free form on the GPU against the oracle:
the shock detector (DGAdapt::update_viscosity_element), the interpolated
viscous flux (LargViscosInterpolation::support_gradu_Lagr_fast, with and
without support overlap) and its volume term, alone and after a Hermite
flux right-hand side as in viscosity_option 3 of
example/02_hyperbolic_05_burgers_adapt.cpp:440-458.  Bit-exact."""
import numpy as np
import pytest

import pyadg
import pyoracle
from common import first_diff, ic_params, init_both, same_grid

pytestmark = pytest.mark.gpu


def circle_ic(radius=1.0):
    return ic_params(3, inside=7 * np.pi / 4, outside=np.pi / 4, radius=radius, center=[0.0, 0.0],
                     scale=[4.0, 4.0], offset=[-2.0, -2.5])


def pair(params, nmax=6, level_init=4):
    print(f"[case] viscosity nmax={nmax}", flush=True)
    g = pyadg.Solution(2, 2, 3, nmax, True, level_init)
    o = pyoracle.Oracle(2, 2, 3, nmax, True, level_init)
    init_both(g, o, params, 1e-3)
    assert same_grid(g, o)
    return g, o


@pytest.mark.parametrize("kappa", [0.0, 0.5])
@pytest.mark.parametrize("overlap", [False, True])
def test_viscous_flux_and_volume_term_bitexact(kappa, overlap):
    g, o = pair(circle_ic())
    n, flags = g.viscosity_elements(kappa)
    m, oflags = o.viscosity_elements(kappa)
    print(f"[viscosity] kappa {kappa}: {n} viscosity elements of {g.num_elem}", flush=True)
    assert n == m and n > 0 and np.array_equal(flags, oflags)
    coef = -1.0 / 64
    g.viscosity_flux(coef, overlap)
    o.viscosity_flux(coef, overlap)
    for d in range(2):
        a, b = g.get(pyadg.FIELD_FUCOE + d), o.get(pyoracle.FIELD_FUCOE + d)
        assert np.array_equal(a, b), f"fucoe[{d}]: " + first_diff(a, b)
        assert np.any(a != 0)
    g.rhs_zero(); o.rhs_zero()
    g.rhs_lagr_vol(); o.rhs_lagr_vol()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), "viscous volume term: " + first_diff(a, b)


def test_hermite_flux_plus_viscosity_rhs_bitexact():
    # the right-hand side of viscosity_option 3: Hermite flux terms, then the Lagrange viscous volume term
    # one right-hand side on discontinuous data (no time stepping)
    g, o = pair(circle_ic(1.05))
    n, _ = g.viscosity_elements(-1.0)
    print(f"[viscosity] hermite case: {n} viscosity elements of {g.num_elem}", flush=True)
    assert n == o.viscosity_elements(-1.0)[0] and n > 0
    for s, fam in ((g, pyadg.INTERP_HERMITE), (o, 1)):
        s.set_interpolation(fam)
        s.eval_up()
        s.flux(0)
        s.fp_to_coe()
        s.rhs_zero()
        s.rhs_lagr_vol()
        s.rhs_lagr_flx()
        s.viscosity_flux(-0.5 / 64, False)
        s.set_interpolation(0)
        s.rhs_lagr_vol()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), first_diff(a, b)


def test_batch_instances_use_their_own_viscosity_elements():
    ics = [circle_ic(0.9), circle_ic(1.15)]
    b = pyadg.Solution(2, 2, 3, 6, True, 4)
    b.init_batch(2, True, 4)
    b.init_adaptive_intp_builtin_batch(ics, 1e-3)
    n, flags = b.viscosity_elements(0.0)
    b.viscosity_flux(-1.0 / 64, True)
    fu = [b.get(pyadg.FIELD_FUCOE + d) for d in range(2)]
    total = 0
    for i, ic in enumerate(ics):
        g, o = pair(ic)
        m, oflags = o.viscosity_elements(0.0)
        first, count = b.instance_range(i)
        assert np.array_equal(flags[first:first + count], oflags)
        total += m
        o.viscosity_flux(-1.0 / 64, True)
        for d in range(2):
            a = fu[d][first * 16:(first + count) * 16]
            assert np.array_equal(a, o.get(pyoracle.FIELD_FUCOE + d)), f"instance {i} fucoe[{d}]"
    assert n == total


def test_viscosity_elements_survive_refinement():
    # the reference holds the viscosity elements by pointer, so the refine of the predict step keeps them;
    # the library keeps them by key across the renumbering that refine makes
    g, o = pair(circle_ic(), nmax=6, level_init=3)
    n, flags = g.viscosity_elements(0.0)
    assert n == o.viscosity_elements(0.0)[0] and n > 0
    before = g.num_elem
    g.refine(1e-5)
    o.refine(1e-5)
    print(f"[viscosity] refine: {before} -> {g.num_elem} elements, {n} viscosity elements", flush=True)
    assert g.num_elem > before and same_grid(g, o)
    g.viscosity_flux(-1.0 / 64, True)
    o.viscosity_flux(-1.0 / 64, True)
    for d in range(2):
        a, b = g.get(pyadg.FIELD_FUCOE + d), o.get(pyoracle.FIELD_FUCOE + d)
        assert np.array_equal(a, b), f"fucoe[{d}]: " + first_diff(a, b)
