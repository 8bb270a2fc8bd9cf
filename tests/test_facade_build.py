"""The C++ facade (reference class names over the C-ABI) builds, drivers written
against it link, and the reference's own hyperbolic example drivers (constant
and variable coefficient, with and without coarse grids) compile and link
unchanged against the facade headers (when the reference tree is present)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "adaptive-multiresolution-dg_amd", "facade")
REF_EXAMPLES = "/root/reference/example"
REF_DRIVERS = ["02_hyperbolic_01_scalar_const_coefficient", "02_hyperbolic_01_scalar_const_coefficient_coarse_grid",
               "02_hyperbolic_02_var_coeff", "02_hyperbolic_02_var_coeff_coarse_grid"]


def test_facade_and_drivers_build():
    subprocess.check_call(["make", "-s", "-C", FACADE])
    for exe in ("burgers_adapt_lagr", "linear_advection"):
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
