"""Step graphs (adg_set_graph_step): a whole fused RK step recorded into a HIP
graph and replayed must give exactly the values of the same steps run
without graphs, and a grid change must make the context record a new graph
(the old one holds the previous grid's pointers and launch shapes)."""
import numpy as np
import pytest

import pyadg
import workloads
from common import first_diff, ic_params, problem

pytestmark = pytest.mark.gpu


def _pair(variant, precision=pyadg.PREC_FP64):
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6))])
    prob = problem(flux_kind=0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=1.0)
    sols = []
    for graph in (0, 1):
        s = pyadg.Solution(2, 2, 3, 6, True, 4)
        s.init_adaptive_intp_builtin(p, 1e-4)
        s.set_kernel_variant(variant)
        s.set_precision(precision)
        s.set_graph_step(graph)
        sols.append(s)
    return sols, prob


@pytest.mark.parametrize("variant,precision", [(-1, pyadg.PREC_FP64), (3, pyadg.PREC_FP64), (7, pyadg.PREC_FP64),
                                               (7, pyadg.PREC_FP32_SPLIT)])
def test_graph_steps_match_and_rerecord_on_grid_change(variant, precision):
    print(f"[case] graph steps variant={variant} precision={precision}", flush=True)
    (a, b), prob = _pair(variant, precision)
    dt = 1e-3

    def steps(n):
        for i in range(n):
            a.rk_step(prob, pyadg.RK3, dt)
            b.rk_step(prob, pyadg.RK3, dt)
            x, y = a.get(pyadg.FIELD_UCOE), b.get(pyadg.FIELD_UCOE)
            assert np.array_equal(x, y), f"step {i}: " + first_diff(x, y)

    # first step on a grid: no graph; then one recording per buffer of the
    # u / u_next pair, then replays
    steps(6)
    assert b.graph_step_count() == (2, 3)
    l0, s0 = a.stage_launch_count()
    l1, s1 = b.stage_launch_count()
    assert s0 == s1, "replayed steps count their stages"
    if variant >= 0:  # (the autotuner of each context may pick other mappings)
        assert l0 == l1, "replayed steps count their launches"
    # a grid change (refinement) must record new graphs
    e0 = b.num_elem
    na, nb = a.refine(1e-5), b.refine(1e-5)
    assert na == nb and na > 0 and b.num_elem > e0
    steps(5)
    assert b.graph_step_count() == (4, 5)
    # and so must a new step size
    dt = 5e-4
    steps(3)
    assert b.graph_step_count() == (6, 5)


def test_graph_off_drops_recordings():
    (a, b), prob = _pair(-1)
    for _ in range(4):
        b.rk_step(prob, pyadg.RK3, 1e-3)
    assert b.graph_step_count() == (2, 1)
    b.set_graph_step(0)
    for _ in range(2):
        b.rk_step(prob, pyadg.RK3, 1e-3)
    assert b.graph_step_count() == (2, 1)


def test_graph_run_hyperbolic_c4_batch():
    """The fused run driver with step graphs on a small C4 batch (per-instance
    steps, adaptive) equals the run without graphs."""
    seeds = [3, 5]
    print(f"[case] graph run c4 seeds={seeds}", flush=True)
    out = []
    for graph in (0, 1):
        ws = [workloads.c4_advection_3d(seed=s, nmax=5, level_init=2) for s in seeds]
        b = workloads.make_batch(ws)
        b.set_kernel_variant(7)
        b.set_graph_step(graph)
        prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
        prob.form, prob.max_steps = 1, 6
        b.run_hyperbolic(prob)
        out.append((b, [b.instance_get(i) for i in range(len(seeds))], b.graph_step_count()))
    for i in range(len(seeds)):
        x, y = out[0][1][i], out[1][1][i]
        assert np.array_equal(x, y), f"instance {i}: " + first_diff(x, y)
    assert out[0][2] == (0, 0)
    print(f"[graph] c4 batch captures/replays={out[1][2]}", flush=True)


def test_graph_fixed_grid_run_replays():
    """A run on a fixed grid (no adaptivity) replays every step after the
    first three."""
    p = ic_params(1, shift=0.1)
    prob = problem(flux_kind=2, flux_param=0.5, alpha=[1.5] * 3, wave_speed=[1.5] * 3, cfl=0.2, t_final=1.0)
    prob.form, prob.adaptive, prob.max_steps = 1, 0, 8
    res = []
    for graph in (0, 1):
        s = pyadg.Solution(3, 1, 2, 4, True, 3)
        s.init_adaptive_intp_builtin(p, 1e-4)
        s.set_kernel_variant(7)
        s.set_graph_step(graph)
        s.run_hyperbolic(prob)
        res.append((s.get(pyadg.FIELD_UCOE), s.graph_step_count()))
    assert np.array_equal(res[0][0], res[1][0]), first_diff(res[0][0], res[1][0])
    assert res[1][1] == (2, 5)
