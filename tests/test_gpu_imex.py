"""The assembled artificial viscosity operator (viscosity options 1 and 2) and
IMEX43 on the GPU against the oracle.

The operator and rhs += (A u) are bit-exact (same 1D integral table, entries
summed in viscosity-element order, rows in increasing column order).  IMEX43's
conjugate gradient sums its dot products in a different order from the
oracle's, so the stage values agree to 1e-10 relative (the solver's tolerance
is 1e-12 relative residual); a batch of instances equals each instance alone
bit for bit (per-instance reductions over the same chunks)."""
import numpy as np
import pytest

import pyadg
import pyoracle
from common import first_diff, ic_params, init_both, oracle_hyperbolic_rhs, problem, same_grid

pytestmark = pytest.mark.gpu


def circle_ic(radius=1.0):
    return ic_params(3, inside=7 * np.pi / 4, outside=np.pi / 4, radius=radius, center=[0.0, 0.0],
                     scale=[4.0, 4.0], offset=[-2.0, -2.5])


def pair(params, nmax=6, level_init=4, dim=2, k=2, eps=1e-3):
    print(f"[case] imex dim={dim} nmax={nmax} level_init={level_init} eps={eps}", flush=True)
    g = pyadg.Solution(dim, k, 3, nmax, True, level_init)
    o = pyoracle.Oracle(dim, k, 3, nmax, True, level_init)
    init_both(g, o, params, eps)
    assert same_grid(g, o)
    return g, o


@pytest.mark.parametrize("grid", ["adaptive", "sparse"])
@pytest.mark.parametrize("option", [1, 2])
def test_operator_and_rhs_bitexact(option, grid):
    # adaptive: the circle's refined grid; sparse: the sparse grid N = 5 without refinement, where the
    # related sets of options 1 and 2 differ (elements finer than w in one dimension meet it)
    if grid == "adaptive":
        g, o = pair(circle_ic())
        kappa = 0.0
    else:
        g, o = pair(circle_ic(), nmax=5, level_init=5, eps=1e10)
        kappa = -100.0
    n = g.viscosity_elements(kappa)[0]
    assert n == o.viscosity_elements(kappa)[0] and n > 0
    blocks = g.viscosity_operator(-1.0 / 64, option)
    nnz = o.viscosity_operator(-1.0 / 64, option)
    print(f"[imex] {grid} grid, option {option}: {n} viscosity elements, {blocks} blocks, {nnz} oracle entries",
          flush=True)
    if grid == "sparse" and option == 2:
        assert blocks < g.viscosity_operator(-1.0 / 64, 1)
        g.viscosity_operator(-1.0 / 64, 2)
    rng = np.random.default_rng(option)
    r0 = rng.standard_normal(g.num_elem * g.K)
    for s in (g, o):
        s.set(pyadg.FIELD_RHS, r0)
        s.rhs_viscosity_operator()
    a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
    assert np.array_equal(a, b), first_diff(a, b)
    assert not np.array_equal(a, r0)


def test_operator_3d_bitexact():
    g, o = pair(ic_params(2, base=0.5, ampl=0.25, bumps=[(0.3, 0.05, (0.3, 0.6, 0.5))]), nmax=3, level_init=3,
                dim=3, k=1, eps=1e10)
    n = g.viscosity_elements(-100.0)[0]
    assert n == o.viscosity_elements(-100.0)[0] and n > 0
    for option in (1, 2):
        g.viscosity_operator(-0.25, option)
        o.viscosity_operator(-0.25, option)
        for s in (g, o):
            s.rhs_zero()
            s.rhs_viscosity_operator()
        a, b = g.get(pyadg.FIELD_RHS), o.get(pyoracle.FIELD_RHS)
        assert np.array_equal(a, b), f"option {option}: " + first_diff(a, b)
        assert np.any(a != 0)


def _prob():
    return problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.05, t_final=1.0)


def _imex_step(s, dt, prob, log):
    s.imex43_init(dt)
    for st in range(5):
        if st >= 2:
            oracle_hyperbolic_rhs(s, prob)
        log.append(s.imex43_stage(st))


@pytest.mark.parametrize("option", [1, 2])
def test_imex43_stages_match_the_oracle(option):
    g, o = pair(circle_ic())
    prob = _prob()
    dt = 0.05 / (2 * 2 ** 6)
    for step in range(2):
        ng = g.viscosity_elements(0.0)[0]
        assert ng == o.viscosity_elements(0.0)[0]
        g.viscosity_operator(-2.0 / 64, option)
        o.viscosity_operator(-2.0 / 64, option)
        g.imex43_init(dt)
        o.imex43_init(dt)
        for st in range(5):
            if st >= 2:
                oracle_hyperbolic_rhs(g, prob)
                oracle_hyperbolic_rhs(o, prob)
            ig = g.imex43_stage(st)
            io = o.imex43_stage(st)
            a, b = g.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
            err = np.max(np.abs(a - b)) / np.max(np.abs(b))
            print(f"[imex] option {option} step {step} stage {st}: CG iterations {ig} / oracle {io}, "
                  f"relative difference {err:.2e}", flush=True)
            assert io >= 0 and abs(ig - io) <= 1
            assert err < 1e-10
            if st < 4:
                assert ig > 0 or ng == 0


def test_imex43_batch_equals_single_solves():
    ics = [circle_ic(0.9), circle_ic(1.15)]
    prob = _prob()
    dt = 0.05 / (2 * 2 ** 6)
    b = pyadg.Solution(2, 2, 3, 6, True, 4)
    b.init_batch(2, True, 4)
    b.init_adaptive_intp_builtin_batch(ics, 1e-3)
    b.viscosity_elements(0.0)
    b.viscosity_operator(-2.0 / 64, 1)
    _imex_step(b, dt, prob, [])
    u = b.get(pyadg.FIELD_UCOE)
    for i, ic in enumerate(ics):
        s = pyadg.Solution(2, 2, 3, 6, True, 4)
        s.init_adaptive_intp_builtin(ic, 1e-3)
        s.viscosity_elements(0.0)
        s.viscosity_operator(-2.0 / 64, 1)
        _imex_step(s, dt, prob, [])
        first, count = b.instance_range(i)
        a = u[first * 9:(first + count) * 9]
        assert np.array_equal(a, s.get(pyadg.FIELD_UCOE)), f"instance {i}: " + first_diff(a, s.get(pyadg.FIELD_UCOE))


def test_operator_is_invalidated_by_a_grid_change():
    g = pyadg.Solution(2, 2, 3, 6, True, 3)
    g.init_adaptive_intp_builtin(circle_ic(), 1e-3)
    g.viscosity_elements(0.0)
    g.viscosity_operator(-1.0 / 64, 1)
    g.rhs_zero()
    g.rhs_viscosity_operator()
    assert g.refine(1e-5) > 0
    with pytest.raises(RuntimeError):
        g.rhs_viscosity_operator()
    with pytest.raises(RuntimeError):
        g.viscosity_operator(-1.0 / 64, 3)
