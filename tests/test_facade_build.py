"""The C++ facade (reference class names over the C-ABI) builds, drivers written
against it link, and the reference's own config-1 example driver compiles and
links unchanged against the facade headers (when the reference tree is present)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "adaptive-multiresolution-dg_amd", "facade")
REF_DRIVER = "/root/reference/example/02_hyperbolic_01_scalar_const_coefficient.cpp"


def test_facade_and_drivers_build():
    subprocess.check_call(["make", "-s", "-C", FACADE])
    for exe in ("burgers_adapt_lagr", "linear_advection"):
        assert os.access(os.path.join(FACADE, exe), os.X_OK)


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference tree not mounted")
def test_reference_driver_links_unchanged():
    subprocess.check_call(["make", "-s", "-C", FACADE, "libadg_facade.a"])
    out = os.path.join(tempfile.mkdtemp(), "ref_c1")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(FACADE, "include"),
                           "-I", os.path.join(ROOT, "include"), REF_DRIVER, os.path.join(FACADE, "libadg_facade.a"),
                           "-L", os.path.join(ROOT, "adaptive-multiresolution-dg_amd"), "-ladg_hip", "-o", out])
    assert os.access(out, os.X_OK)
