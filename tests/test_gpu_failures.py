"""Failure detection, forced from the host: every kernel exits normally and
the device error flag reports the condition through the C-ABI.

- a NaN planted in the coefficients makes the next RK stage non-finite:
  the run drivers and adg_sync return ADG_E_NONFINITE (the reference drivers
  only notice a blow-up through their L2 check,
  example/02_hyperbolic_02_var_coeff_coarse_grid.cpp:389-396);
- a grid with a hole (an element whose parent is missing) makes the surplus
  transform miss an ancestor: adg_fp_to_coe returns ADG_E_ANCESTOR (the
  reference asserts on the failed hash lookup, source/Interplation.cpp:1357-1359).
"""
import numpy as np
import pytest

import pyadg
from common import ic_params, problem

pytestmark = pytest.mark.gpu

ADG_E_ANCESTOR, ADG_E_NONFINITE = -3, -4


def _burgers(nmax=5):
    g = pyadg.Solution(2, 2, 3, nmax, True, 3)
    g.init_adaptive_intp_builtin(ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.4, 0.6))]), 1e-4)
    return g


@pytest.mark.parametrize("form", [0, 1])
def test_nan_in_solution_is_reported_by_the_run(form):
    print(f"[case] NaN injection, run form {form}", flush=True)
    g = _burgers()
    u = g.get(pyadg.FIELD_UCOE)
    u[len(u) // 2] = np.nan
    g.set(pyadg.FIELD_UCOE, u)
    prob = problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=0.1, adaptive=0,
                   max_steps=3, form=form)
    with pytest.raises(pyadg.AdgError) as ei:
        g.run_hyperbolic(prob)
    assert ei.value.code == ADG_E_NONFINITE
    # the flag was cleared: a clean state steps again
    g2 = _burgers()
    g2.run_hyperbolic(prob)


def test_nan_in_stage_is_reported_by_sync():
    print("[case] NaN injection, adg_stage + adg_sync", flush=True)
    g = _burgers()
    u = g.get(pyadg.FIELD_UCOE)
    u[3] = np.inf
    g.set(pyadg.FIELD_UCOE, u)
    prob = problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=0.1)
    g.rk_init()
    g.stage(prob, pyadg.RK3, 0, 1e-3)
    with pytest.raises(pyadg.AdgError) as ei:
        g.sync()
    assert ei.value.code == ADG_E_NONFINITE
    g.sync()  # reported once, then cleared


def test_hole_in_grid_is_reported_by_the_surplus_transform():
    print("[case] grid with a hole, adg_fp_to_coe", flush=True)
    g = pyadg.Solution(2, 2, 3, 5, True, 3)
    lv, sp, _ = g.grid()
    # drop the level-(1, 0) element: its children at level (2, 0) keep existing
    keep = ~((lv[:, 0] == 1) & (lv[:, 1] == 0))
    assert keep.sum() == len(keep) - 1
    g.set_grid(lv[keep], sp[keep])
    g.set(pyadg.FIELD_UCOE, np.ones(g.num_elem * g.K))
    g.eval_up()
    g.flux(pyadg.FLUX_BURGERS)
    with pytest.raises(pyadg.AdgError) as ei:
        g.fp_to_coe()
    assert ei.value.code == ADG_E_ANCESTOR
