"""Batches of independent solves in one context, against the CPU oracle.

Each instance of a batch must end with exactly the element set and the
coefficients (bit for bit) that the oracle produces for that problem alone:
instances share every launch but no fiber, and each advances with its own
time step size and thresholds.  Instances that need fewer steps leave the
batch early and must be put back unchanged.
"""
import ctypes

import numpy as np
import pytest

import pyadg
import pyoracle
import workloads
from common import first_diff, oracle_run_hyperbolic

pytestmark = pytest.mark.gpu


def _oracle_single(w, eps, eta, form_prob):
    o = pyoracle.Oracle(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    o._keep.append(w.ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(w.ic), ctypes.c_void_p), w.init_eps)
    p = form_prob
    p.eps, p.eta = eps, eta
    steps, dof_stages, _ = oracle_run_hyperbolic(o, p)
    lv, sp, hk = o.grid()
    return steps, dof_stages, lv, sp, hk, o.get(pyoracle.FIELD_UCOE)


def _small_c2(seed, t_final):
    return workloads.c2_burgers(seed=seed, nmax=6, level_init=4, t_final=t_final)


@pytest.mark.parametrize("form", [0, 1])
def test_batch_matches_single_solves(form):
    seeds = [3, 1, 7]
    ws = [_small_c2(s, 0.03) for s in seeds]
    print(f"[case] batch seeds={seeds} nmax=6 form={form}", flush=True)
    b = workloads.make_batch(ws)
    assert b.num_instances == 3
    prob = ws[0].problem
    prob.form = form
    st = b.run_hyperbolic(prob)
    total_dof_stages, max_steps = 0, 0
    for i, w in enumerate(ws):
        w.problem.form = form
        steps, dofs, lv, sp, hk, u = _oracle_single(w, w.problem.eps, w.problem.eta, w.problem)
        total_dof_stages += dofs
        max_steps = max(max_steps, steps)
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv) and np.array_equal(bsp, sp), \
            f"instance {i}: element sets differ"
        bu = b.instance_get(i)
        assert np.array_equal(bu, u), f"instance {i}: " + first_diff(bu, u)
    assert st.steps == max_steps
    assert st.dof_stages == total_dof_stages


def test_batch_per_instance_thresholds():
    seeds = [2, 5]
    ws = [_small_c2(s, 0.02) for s in seeds]
    eps = [1e-4, 5e-4]
    eta = [1e-5, 2e-5]
    print(f"[case] batch thresholds seeds={seeds} eps={eps}", flush=True)
    b = workloads.make_batch(ws)
    b.set_instance_thresholds(eps, eta)
    prob = ws[0].problem
    prob.form = 1
    b.run_hyperbolic(prob)
    for i, w in enumerate(ws):
        w.problem.form = 1
        _, _, lv, sp, hk, u = _oracle_single(w, eps[i], eta[i], w.problem)
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv), f"instance {i}: element sets differ"
        bu = b.instance_get(i)
        assert np.array_equal(bu, u), f"instance {i}: " + first_diff(bu, u)


def test_large_batch_matches_single_gpu_solves():
    # more instances than the serial threshold of the host pool: table builds run on several threads
    seeds = list(range(11, 23))
    ws = [workloads.c2_burgers(seed=s, nmax=5, level_init=3, t_final=0.02) for s in seeds]
    print(f"[case] batch of {len(seeds)} instances, nmax=5", flush=True)
    b = workloads.make_batch(ws)
    prob = ws[0].problem
    prob.form = 1
    b.run_hyperbolic(prob)
    for i, w in enumerate(ws):
        s = workloads.make_solution(w)
        w.problem.form = 1
        s.run_hyperbolic(w.problem)
        lv, sp, hk = s.grid()
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv), f"instance {i}: element sets differ"
        a, c = b.instance_get(i), s.get(pyadg.FIELD_UCOE)
        assert np.array_equal(a, c), f"instance {i}: " + first_diff(a, c)
