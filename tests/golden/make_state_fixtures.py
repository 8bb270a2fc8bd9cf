"""Writes the golden state files of tests/golden from the CPU oracle.

Each case is a short adaptive Burgers run (Lagrange or Hermite flux
interpolation) on a small grid; the files hold the element set and the Alpert
coefficients after the initialisation and after the run.  The GPU must
reproduce them byte for byte; the CPU suite re-runs the oracle and checks that
it still produces them.

    python tests/golden/make_state_fixtures.py
"""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("tests", "oracle", "adaptive-multiresolution-dg_amd"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import pyadg  # noqa: E402
import pyoracle  # noqa: E402
from common import ic_params, problem  # noqa: E402
from runs import run_hyperbolic  # noqa: E402
from state_file import from_oracle  # noqa: E402

CASES = {
    "burgers_lagr_2d_n5": dict(family=0),
    "burgers_herm_2d_n5": dict(family=1),
}


def case_ic():
    return ic_params(2, base=0.5, ampl=0.25, bumps=[(0.15, 0.08, (0.3, 0.6)), (0.1, 0.1, (0.7, 0.3))])


def case_problem():
    return problem(flux_kind=0, flux_param=0.0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2, t_final=0.02,
                   adaptive=1, eps=1e-4, eta=1e-5, max_steps=4)


def run_case(name):
    """(initial state bytes, final state bytes) of a case, from the oracle."""
    fam = CASES[name]["family"]
    o = pyoracle.Oracle(2, 2, 3, 5, True, 3)
    ic = case_ic()
    o._keep.append(ic)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(ic), ctypes.c_void_p), 1e-4)
    o.set_interpolation(fam)
    first = from_oracle(o)
    run_hyperbolic(o, case_problem())
    return first, from_oracle(o)


def path(name, which):
    return os.path.join(HERE, f"{name}_{which}.adgstate")


if __name__ == "__main__":
    for name in CASES:
        a, b = run_case(name)
        for which, data in (("t0", a), ("final", b)):
            with open(path(name, which), "wb") as f:
                f.write(data)
            print(path(name, which), len(data), "bytes")
