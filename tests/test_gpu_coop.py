"""The persistent form of the fused stage: one cooperative launch per RK stage
(adg_set_stage_launch(1)) must give exactly the per-sweep path's values (FP64)
and exactly the per-phase launches' values (FP32), with one launch per stage.

The persistent launch is switched off: runs of these cases on the device
stopped finishing (a 4D case, then a 3D case that had finished in an
earlier run with the same inputs) and the cause is not found.  The API
refuses mode 1 and the cases below stay skipped."""
import numpy as np
import pytest

import pyadg
import pyoracle
import workloads
from common import first_diff, ic_params, problem, same_grid
from test_gpu_parity import adapted_pair
from test_gpu_c4 import _oracle

pytestmark = pytest.mark.gpu


def _case(kind, dim):
    if kind == 0:
        p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6))])
        prob = problem(flux_kind=0, alpha=[1.2] * dim, wave_speed=[1.0] * dim, cfl=0.2, t_final=1.0)
    elif kind == 1:
        p = ic_params(3, inside=7 * np.pi / 4, outside=np.pi / 4, radius=1.0, center=[0.0, 0.0], scale=[4.0, 4.0],
                      offset=[-2.0, -2.5])
        prob = problem(flux_kind=1, flux_param=0.25, alpha=[0.25] * dim, wave_speed=[0.25] * dim, cfl=0.2,
                       t_final=1.0)
    else:
        p = ic_params(1, shift=0.1)
        prob = problem(flux_kind=2, flux_param=0.5, alpha=[1.5] * dim, wave_speed=[1.5] * dim, cfl=0.2, t_final=1.0)
    return p, prob


OFF = pytest.mark.skip(reason="persistent cooperative stage switched off: its launch did not always finish on the "
                              "device (cause not found)")


def test_persistent_stage_is_refused():
    s = pyadg.Solution(2, 1, 2, 4, True, 2)
    with pytest.raises(pyadg.AdgError, match="switched off"):
        s.set_stage_launch(1)
    s.set_stage_launch(0)


@OFF
@pytest.mark.parametrize("dim,k,m,nmax,level_init,kind", [(2, 2, 3, 6, 4, 0), (2, 2, 3, 6, 4, 1), (3, 1, 2, 4, 3, 2),
                                                          (4, 1, 2, 4, 4, 2), (2, 3, 4, 5, 4, 2)])
def test_persistent_stage_matches_per_sweep_path(dim, k, m, nmax, level_init, kind):
    print(f"[case] persistent stage dim={dim} k={k} m={m} kind={kind}", flush=True)
    p, prob = _case(kind, dim)
    a, _ = adapted_pair(dim, k, m, nmax, level_init, p, 1e-4)
    b = pyadg.Solution(dim, k, m, nmax, True, level_init)
    b.init_adaptive_intp_builtin(p, 1e-4)
    b.set_stage_launch(1)
    assert same_grid(a, b)
    l0, s0 = b.stage_launch_count()
    a.rk_init(); b.rk_init()
    for st in range(3):
        a.eval_up(); a.flux(prob.flux_kind, (prob.flux_param,)); a.fp_to_coe(); a.rhs_zero(); a.rhs_lagr_vol()
        a.rhs_lagr_flx(); a.rhs_penalty([prob.alpha[d] for d in range(dim)]); a.rk_stage(pyadg.RK3, st, 1e-3)
        b.stage(prob, pyadg.RK3, st, 1e-3)
        for f in (pyadg.FIELD_UP, pyadg.FIELD_RHS, pyadg.FIELD_UCOE):
            x, y = a.get(f), b.get(f)
            assert np.array_equal(x, y), f"stage {st} field {f}: " + first_diff(x, y)
    l1, s1 = b.stage_launch_count()
    assert (l1 - l0, s1 - s0) == (3, 3), "one launch per stage"


@OFF
def test_persistent_fp32_equals_tile_launches():
    p, prob = _case(0, 2)
    outs = []
    for mode in (0, 1):
        s = pyadg.Solution(2, 2, 3, 6, True, 4)
        s.init_adaptive_intp_builtin(p, 1e-4)
        s.set_precision(pyadg.PREC_FP32_SPLIT)
        s.set_kernel_variant(7)
        s.set_stage_launch(mode)
        s.rk_init()
        for st in range(3):
            s.stage(prob, pyadg.RK3, st, 1e-3)
        outs.append(s.get(pyadg.FIELD_UCOE))
    assert np.array_equal(outs[0], outs[1]), first_diff(outs[0], outs[1])


@OFF
def test_persistent_c4_batch_matches_oracle():
    seeds = [2, 6, 9]
    ws = [workloads.c4_advection_3d(seed=s, nmax=5, level_init=2) for s in seeds]
    print(f"[case] persistent c4 batch seeds={seeds}", flush=True)
    b = workloads.make_batch(ws)
    b.set_stage_launch(1)
    prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
    prob.form, prob.max_steps = 1, 3
    b.run_hyperbolic(prob)
    for i, w in enumerate(ws):
        o, _, _ = _oracle(w, 3)
        lv, sp, hk = o.grid()
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv), f"instance {i}: element sets differ"
        a, c = b.instance_get(i), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, c), f"instance {i}: " + first_diff(a, c)
