"""Refine and coarsen on the device (adg_set_device_adapt).

The device path computes the threshold tests, children, hole closure, hash
keys, the sorted element table and its fiber tables on the GPU.  Its element
sets, their order, the fiber tables and the coefficients must equal the host
path's and the CPU oracle's bit for bit; adg_grid_check compares every device
table entry with a host build of the same grid.
"""
import ctypes
import math

import numpy as np
import pytest

import pyadg
import pyoracle
import workloads
from common import first_diff, ic_params, oracle_run_hyperbolic

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    ga, gb = a.grid(), b.grid()
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y), f"{what}: element tables differ ({len(ga[2])} vs {len(gb[2])} elements)"
    x, y = a.get(pyadg.FIELD_UCOE), b.get(pyadg.FIELD_UCOE)
    assert np.array_equal(x, y), f"{what}: " + first_diff(x, y)


def _pair(dim, k, m, nmax, level_init, params, eps):
    out = []
    for dev in (0, 1):
        s = pyadg.Solution(dim, k, m, nmax, True, level_init)
        s.init_adaptive_intp_builtin(params, eps)
        s.set_device_adapt(dev)
        out.append(s)
    return out


def _bump(dim):
    return ic_params(2, base=0.5, ampl=0.25,
                     bumps=[(0.4, 0.07, tuple([0.3, 0.6, 0.45, 0.55][:dim])),
                            (0.2, 0.05, tuple([0.7, 0.25, 0.6, 0.35][:dim]))])


@pytest.mark.parametrize("dim,k,m,nmax", [(1, 2, 3, 9), (2, 2, 3, 7), (3, 2, 3, 5), (4, 1, 2, 4)])
def test_refine_and_coarsen_calls_match_host(dim, k, m, nmax):
    print(f"[case] device adapt calls dim={dim} k={k} m={m} nmax={nmax}", flush=True)
    a, b = _pair(dim, k, m, nmax, 2, _bump(dim), 1e-3)
    _same(a, b, "initial")
    assert b.grid_check() == 0
    for eps, eta in [(1e-4, 1e-5), (2e-5, 1e-6), (1e-5, 3e-4)]:
        na, nb = a.refine(eps), b.refine(eps)
        assert na == nb
        assert b.grid_check() == 0, f"refine {eps}: device tables differ from the host build"
        _same(a, b, f"refine {eps}")
        ca, cb = a.coarsen(eta), b.coarsen(eta)
        assert ca == cb
        assert b.grid_check() == 0, f"coarsen {eta}: device tables differ from the host build"
        _same(a, b, f"coarsen {eta}")
    dev, host = b.device_adapt_stats()
    print(f"[device adapt] calls on the device {dev}, handed to the host {host}", flush=True)
    assert dev == 6 and host == 0
    assert a.device_adapt_stats() == (0, 0)


def test_norm_at_the_threshold_goes_to_the_host():
    """A norm within a few ulps of eps is decided by the reference's pow on
    the host; the result is the host path's."""
    print("[case] device adapt band fallback", flush=True)
    a, b = _pair(2, 2, 3, 6, 2, _bump(2), 1e-3)
    s = b.indicator_sumsq()
    e = int(np.argmax(s > 1e-12))
    eps = math.sqrt(s[e])
    na, nb = a.refine(eps), b.refine(eps)
    assert na == nb
    _same(a, b, "refine at a norm")
    assert b.grid_check() == 0
    dev, host = b.device_adapt_stats()
    assert host == 1 and dev == 0


def _oracle(w, max_steps):
    o = pyoracle.Oracle(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    o._keep.append(w.ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(w.ic), ctypes.c_void_p), w.init_eps)
    p = w.oracle_problem()
    p.max_steps = max_steps
    oracle_run_hyperbolic(o, p)
    return o


def test_c4_batch_run_with_device_adapt_matches_oracle():
    seeds = [1, 2, 3, 4]
    ws = [workloads.c4_advection_3d(seed=s, nmax=5, level_init=2) for s in seeds]
    print(f"[case] c4 batch device adapt seeds={seeds} nmax=5", flush=True)
    b = workloads.make_batch(ws)
    b.set_device_adapt(1)
    prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
    prob.form, prob.max_steps = 1, 4
    b.run_hyperbolic(prob)
    dev, host = b.device_adapt_stats()
    print(f"[device adapt] calls on the device {dev}, handed to the host {host}", flush=True)
    assert dev >= 8
    assert b.grid_check() == 0
    for i, w in enumerate(ws):
        o = _oracle(w, 4)
        lv, sp, hk = o.grid()
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv) and np.array_equal(bsp, sp), \
            f"instance {i}: element sets differ ({len(bhk)} vs {len(hk)})"
        x, y = b.instance_get(i), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(x, y), f"instance {i}: " + first_diff(x, y)


@pytest.mark.parametrize("config", ["c2", "c5"])
def test_batch_run_with_device_adapt_matches_host(config):
    """Whole adaptive runs (instances that finish early leave the batch and
    are put back at the end) give the host path's tables and values."""
    if config == "c2":
        ws = [workloads.c2_burgers(seed=s, nmax=6, level_init=4, t_final=0.03) for s in (3, 1, 7)]
    else:
        ws = [workloads.c5_rotating(seed=s, nmax=6, level_init=4, t_final=0.01) for s in (2, 5)]
    print(f"[case] {config} batch run device adapt vs host", flush=True)
    res = []
    for dev in (0, 1):
        b = workloads.make_batch(ws)
        b.set_device_adapt(dev)
        prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
        prob.form = 1
        st = b.run_hyperbolic(prob)
        res.append((b, st))
    (a, sa), (b, sb) = res
    assert (sa.steps, sa.dof_stages, sa.max_elem) == (sb.steps, sb.dof_stages, sb.max_elem)
    for i in range(len(ws)):
        ga, gb = a.instance_grid(i), b.instance_grid(i)
        for x, y in zip(ga, gb):
            assert np.array_equal(x, y), f"instance {i}: element sets differ"
        x, y = a.instance_get(i), b.instance_get(i)
        assert np.array_equal(x, y), f"instance {i}: " + first_diff(x, y)
    dev, host = b.device_adapt_stats()
    print(f"[device adapt] calls on the device {dev}, handed to the host {host}", flush=True)
    assert dev > 0
    assert b.grid_check() == 0


def test_refinement_larger_than_the_new_element_list():
    """A refinement that adds more elements than the first new-element list
    holds (a quarter of the table, at least 2^16) runs again on the device
    with a larger list and still equals the host path."""
    print("[case] device adapt large refinement", flush=True)
    sols = []
    for dev in (0, 1):
        s = pyadg.Solution(2, 1, 1, 9, False, 6)
        s.init_batch(12, sparse=False, level_init=6)  # 8,192 new elements per instance
        s.set_device_adapt(dev)
        rng = np.random.default_rng(11)
        s.set(pyadg.FIELD_UCOE, rng.uniform(0.5, 1.0, s.num_elem * s.K))
        sols.append(s)
    a, b = sols
    e0 = b.num_elem
    na, nb = a.refine(1e-12), b.refine(1e-12)
    print(f"[device adapt] {e0} elements, {nb} added", flush=True)
    assert na == nb and nb > max(e0 // 4, 65536)
    _same(a, b, "large refinement")
    assert b.grid_check() == 0
    # the new elements carry zero coefficients: coarsening removes them again,
    # leaves first, in several passes over a table of more than 2^17 elements
    ca, cb = a.coarsen(1e-3), b.coarsen(1e-3)
    print(f"[device adapt] {cb} removed", flush=True)
    assert ca == cb == nb
    _same(a, b, "large coarsening")
    assert b.grid_check() == 0
    assert b.device_adapt_stats() == (2, 0)


def test_restore_and_rerun_with_device_tables():
    """Restoring a snapshot after a run with device adapt rebuilds the saved
    grid on the device (fields as saved), and a second run repeats the first
    exactly."""
    ws = [workloads.c2_burgers(seed=s, nmax=6, level_init=4, t_final=0.02) for s in (4, 9)]
    print("[case] device adapt restore and rerun", flush=True)
    b = workloads.make_batch(ws)
    b.set_device_adapt(1)
    b.save_state(0)
    start = [(b.instance_grid(i), b.instance_get(i)) for i in range(len(ws))]
    prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
    prob.form = 1
    runs = []
    for r in range(2):
        if r:
            b.restore_state(0)
            assert b.grid_check() == 0
            for i, (g, u) in enumerate(start):
                for x, y in zip(b.instance_grid(i), g):
                    assert np.array_equal(x, y), f"instance {i}: restored grid differs"
                assert np.array_equal(b.instance_get(i), u), f"instance {i}: restored coefficients differ"
        st = b.run_hyperbolic(prob)
        runs.append(((st.steps, st.dof_stages), [(b.instance_grid(i), b.instance_get(i)) for i in range(len(ws))]))
    assert runs[0][0] == runs[1][0]
    for i in range(len(ws)):
        (g1, u1), (g2, u2) = runs[0][1][i], runs[1][1][i]
        for x, y in zip(g1, g2):
            assert np.array_equal(x, y), f"instance {i}: second run's grid differs"
        assert np.array_equal(u1, u2), f"instance {i}: " + first_diff(u1, u2)


@pytest.mark.parametrize("mode", ["current", "learned"])
def test_proposer_with_deferred_judge_on_the_device_matches_host(mode):
    """Proposals, step and deferring judge on the device (refine modes with
    judge_defer = 1) give the host path's grids, values and counts."""
    ws = [workloads.c2_burgers(seed=s, nmax=6, level_init=4, t_final=0.02) for s in (3, 8)]
    print(f"[case] device adapt proposer mode={mode}", flush=True)
    res = []
    for dev in (0, 1):
        b = workloads.make_batch(ws)
        b.set_device_adapt(dev)
        if mode == "learned":
            b.proposer_init(7)  # an untrained network proposes something; both paths must agree
            b.set_refine_mode(pyadg.REFINE_LEARNED, 0.5)
        else:
            b.set_refine_mode(pyadg.REFINE_CURRENT, 0.5)
        b.set_judge_defer(1)
        prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
        prob.form = 1
        st = b.run_hyperbolic(prob)
        res.append((b, (st.steps, st.dof_stages, st.max_elem), b.proposer_stats()))
    (a, sa, pa), (b, sb, pb) = res
    print(f"[device adapt] proposed / redone / deferred: host {pa}, device {pb}", flush=True)
    assert sa == sb and pa == pb
    assert pb[0] > 0
    for i in range(len(ws)):
        for x, y in zip(a.instance_grid(i), b.instance_grid(i)):
            assert np.array_equal(x, y), f"instance {i}: element sets differ"
        x, y = a.instance_get(i), b.instance_get(i)
        assert np.array_equal(x, y), f"instance {i}: " + first_diff(x, y)
    dev, host = b.device_adapt_stats()
    assert dev > 0 and host == 0
    assert b.grid_check() == 0


def test_deferred_judge_norm_at_eps_decided_on_host():
    """The device proposer step sends norms within a few ulps of eps to the
    host element by element (the reference's pow decides them): with eps
    equal to an element's norm, the run equals the host path's."""
    w = workloads.c2_burgers(seed=5, nmax=6, level_init=4, t_final=0.02)
    print("[case] device adapt proposer, eps at a norm", flush=True)
    res = []
    for dev in (0, 1):
        s = workloads.make_solution(w)
        s.set_device_adapt(dev)
        sumsq = s.indicator_sumsq()
        eps = math.sqrt(float(sumsq[int(np.argmax(sumsq))]))
        s.set_refine_mode(pyadg.REFINE_CURRENT, 0.5)
        s.set_judge_defer(1)
        prob = pyadg.Problem.from_buffer_copy(w.problem)
        prob.form, prob.eps, prob.eta, prob.max_steps = 1, eps, eps / 10, 3
        st = s.run_hyperbolic(prob)
        res.append((s, (st.steps, st.dof_stages), s.proposer_stats()))
    (a, sa, pa), (b, sb, pb) = res
    assert sa == sb and pa == pb
    _same(a, b, "eps at a norm")
    assert b.grid_check() == 0
