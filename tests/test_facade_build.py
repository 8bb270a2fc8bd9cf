"""The C++ facade (reference class names over the C-ABI) builds, drivers written
against it link, and the reference's own hyperbolic example drivers (constant
and variable coefficient, with and without coarse grids) compile and link
unchanged against the facade headers (when the reference tree is present)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "adaptive-multiresolution-dg_amd", "facade")
REF_EXAMPLES = "/root/reference/example"
REF_DRIVERS = ["02_hyperbolic_01_scalar_const_coefficient", "02_hyperbolic_01_scalar_const_coefficient_coarse_grid",
               "02_hyperbolic_02_var_coeff", "02_hyperbolic_02_var_coeff_coarse_grid",
               "02_hyperbolic_05_burgers_adapt", "02_hyperbolic_07_kpp"]


def test_facade_and_drivers_build():
    subprocess.check_call(["make", "-s", "-C", FACADE])
    for exe in ("burgers_adapt_lagr", "burgers_adapt_herm", "linear_advection", "var_coeff_coarse_grid"):
        assert os.access(os.path.join(FACADE, exe), os.X_OK)


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference tree not mounted")
@pytest.mark.parametrize("driver", REF_DRIVERS)
def test_reference_driver_links_unchanged(driver):
    subprocess.check_call(["make", "-s", "-C", FACADE, "libadg_facade.a"])
    out = os.path.join(tempfile.mkdtemp(), driver)
    src = os.path.join(REF_EXAMPLES, driver + ".cpp")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(FACADE, "include"),
                           "-I", os.path.join(ROOT, "include"), src, os.path.join(FACADE, "libadg_facade.a"),
                           "-L", os.path.join(ROOT, "adaptive-multiresolution-dg_amd"), "-ladg_hip", "-o", out])
    assert os.access(out, os.X_OK)


BURGERS_HARNESS = r"""
#include "ExactSolution.h"
int main() {
  BurgersExact b(0.3, 0.8, 0.1);
  std::cout.precision(17);
  const double pts[4][2] = {{0.1, 0.2}, {0.45, 0.05}, {0.7, 0.9}, {0.25, 0.25}};
  for (double t : {0.0, 0.02, 0.05})
    for (auto& p : pts) std::cout << p[0] << " " << p[1] << " " << t << " " << b.exact_2d(p[0], p[1], t) << "\n";
  return 0;
}
"""


def test_burgers_exact_solves_the_characteristic_equation(tmp_path):
    # BurgersExact::exact_2d against scipy's root of u = a + b sin(2 pi (x + y - 2 u t + c))
    from scipy.optimize import brentq
    subprocess.check_call(["make", "-s", "-C", FACADE, "libadg_facade.a"])
    src = tmp_path / "h.cpp"
    src.write_text(BURGERS_HARNESS)
    exe = str(tmp_path / "h")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(FACADE, "include"), "-I",
                           os.path.join(ROOT, "include"), str(src), os.path.join(FACADE, "libadg_facade.a"), "-L",
                           os.path.join(ROOT, "adaptive-multiresolution-dg_amd"), "-ladg_hip",
                           "-Wl,-rpath," + os.path.join(ROOT, "adaptive-multiresolution-dg_amd"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    a, b, c = 0.3, 0.8, 0.1
    n = 0
    for line in out:
        if not line.strip():
            continue
        x, y, t, u = map(float, line.split())
        g = lambda v: v - a - b * np.sin(2 * np.pi * (x + y - 2 * v * t + c))  # noqa: E731
        want = brentq(g, a - b - 1e-9, a + b + 1e-9, xtol=1e-15)
        assert abs(u - want) < 1e-12, (x, y, t, u, want)
        n += 1
    assert n == 12
