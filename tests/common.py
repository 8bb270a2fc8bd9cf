"""Shared helpers for the tests: problem definitions and reference-style loops
run on the oracle, mirroring the library's run drivers step by step."""
import ctypes

import numpy as np

import pyadg
import pyoracle

PI = np.pi


def ic_params(kind, **kw):
    p = pyadg.ICParams()
    p.kind = kind
    for name, val in kw.items():
        if name == "bumps":
            p.nbump = len(val)
            for b, (amp, width, center) in enumerate(val):
                p.bump_amp[b] = amp
                p.bump_width[b] = width
                for d, c in enumerate(center):
                    p.bump_center[b][d] = c
        elif name in ("center", "scale", "offset"):
            arr = getattr(p, name)
            for d, c in enumerate(val):
                arr[d] = c
        else:
            setattr(p, name, val)
    return p


def init_both(gpu, orc, params, eps):
    """Initialise a GPU solution and an oracle with the library's built-in
    point function so both evaluate bit-identical point values."""
    gpu.init_adaptive_intp_builtin(params, eps)
    fn = pyadg.ic_eval_ptr()
    orc._keep.append(params)
    orc.init_adaptive_intp_c(fn, ctypes.cast(ctypes.pointer(params), ctypes.c_void_p), eps)


from runs import hyperbolic_rhs as oracle_hyperbolic_rhs  # noqa: E402,F401
from runs import run_hyperbolic as oracle_run_hyperbolic  # noqa: E402,F401
from runs import run_linear as oracle_run_linear  # noqa: E402,F401


def problem(**kw):
    p = pyadg.Problem()
    for name, val in kw.items():
        if name in ("alpha", "wave_speed"):
            arr = getattr(p, name)
            for d, c in enumerate(val):
                arr[d] = c
        else:
            setattr(p, name, val)
    return p


def same_grid(a, b):
    la, sa, ka = a.grid()
    lb, sb, kb = b.grid()
    return np.array_equal(la, lb) and np.array_equal(sa, sb) and np.array_equal(ka, kb)


def first_diff(a, b):
    idx = np.nonzero(a != b)[0]
    if idx.size == 0:
        return "identical"
    i = idx[0]
    return f"{idx.size} differ; first at {i}: {a[i]!r} vs {b[i]!r}; max abs {np.max(np.abs(a - b)):.3e}"
