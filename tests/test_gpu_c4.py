"""The batching configuration (3D variable-coefficient advection, one
instance per seeded problem, per-instance speeds, LF constants and
thresholds) against the CPU oracle, instance by instance.

Every instance of a batch must end with the element set and the
coefficients, bit for bit, that the oracle computes for that problem alone
(flux kind 3: a_d(x) = c_d (1 + s sin(2 pi x_{d+1})) u, the instance's c_d).
The full configuration is NMAX 7; the oracle restates the reference's data
structures and is run here at NMAX 4 and 5 for a few time steps.
"""
import ctypes

import numpy as np
import pytest

import pyadg
import pyoracle
import workloads
from common import first_diff, oracle_run_hyperbolic

pytestmark = pytest.mark.gpu


def _oracle(w, max_steps):
    o = pyoracle.Oracle(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    o._keep.append(w.ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(w.ic), ctypes.c_void_p), w.init_eps)
    p = w.oracle_problem()
    p.max_steps = max_steps
    steps, dof_stages, _ = oracle_run_hyperbolic(o, p)
    return o, steps, dof_stages


@pytest.mark.parametrize("form,variant", [(0, -1), (1, -1), (1, 1), (1, 5)])
def test_c4_batch_matches_oracle_singles(form, variant):
    seeds = [1, 2, 3, 4]
    ws = [workloads.c4_advection_3d(seed=s, nmax=4, level_init=2) for s in seeds]
    print(f"[case] c4 batch seeds={seeds} nmax=4 form={form} variant={variant}", flush=True)
    b = workloads.make_batch(ws)
    b.set_kernel_variant(variant)
    prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
    prob.form, prob.max_steps = form, 3
    st = b.run_hyperbolic(prob)
    tot = 0
    for i, w in enumerate(ws):
        o, steps, dofs = _oracle(w, 3)
        tot += dofs
        lv, sp, hk = o.grid()
        blv, bsp, bhk = b.instance_grid(i)
        assert np.array_equal(bhk, hk) and np.array_equal(blv, lv) and np.array_equal(bsp, sp), \
            f"instance {i}: element sets differ ({len(bhk)} vs {len(hk)})"
        a, c = b.instance_get(i), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, c), f"instance {i}: " + first_diff(a, c)
    assert st.dof_stages == tot


def test_c4_single_instance_plain_and_fused_match_oracle():
    w = workloads.c4_advection_3d(seed=7, nmax=5, level_init=2)
    print("[case] c4 single seed=7 nmax=5", flush=True)
    o, steps, _ = _oracle(w, 2)
    for form in (0, 1):
        s = workloads.make_solution(w)
        prob = pyadg.Problem.from_buffer_copy(w.problem)
        prob.form, prob.max_steps = form, 2
        st = s.run_hyperbolic(prob)
        assert st.steps == steps
        lv, sp, hk = o.grid()
        glv, gsp, ghk = s.grid()
        assert np.array_equal(ghk, hk) and np.array_equal(glv, lv), f"form {form}: element sets differ"
        a, c = s.get(pyadg.FIELD_UCOE), o.get(pyoracle.FIELD_UCOE)
        assert np.array_equal(a, c), f"form {form}: " + first_diff(a, c)


def test_c4_batch_mixed_class_mappings_match_oracle():
    """Fixed mappings per kernel class (the bench's 3,3,7,7 and a mix that
    puts every kernel family in one stage) give the oracle's values, and the
    autotuner's decisions are reported per class."""
    seeds = [4, 8]
    ws = [workloads.c4_advection_3d(seed=s, nmax=5, level_init=2) for s in seeds]
    print(f"[case] c4 class mappings seeds={seeds}", flush=True)
    refs = [_oracle(w, 3)[0] for w in ws]
    for mappings in ([3, 3, 7, 7], [0, 5, 9, 1], [-1, 4, -1, 8]):
        b = workloads.make_batch(ws)
        b.set_class_variants(mappings)
        prob = pyadg.Problem.from_buffer_copy(ws[0].problem)
        prob.form, prob.max_steps = 1, 3
        b.run_hyperbolic(prob)
        for i, o in enumerate(refs):
            a, c = b.instance_get(i), o.get(pyoracle.FIELD_UCOE)
            assert np.array_equal(a, c), f"mappings {mappings} instance {i}: " + first_diff(a, c)
        tuned = b.kernel_tuning()
        assert all(cls in (0, 2) for cls, _ in tuned), tuned  # only the classes left to the autotuner
        assert (-1 not in mappings) == (not tuned)
