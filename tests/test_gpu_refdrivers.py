"""The reference's own example drivers, compiled unchanged against the facade
headers (facade/ref_* , built where the reference tree is present), run on the
GPU: they must finish their time loop and print finite errors / norms."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "adaptive-multiresolution-dg_amd", "facade")

pytestmark = pytest.mark.gpu

CASES = [
    ("02_hyperbolic_02_var_coeff", ["-NM", "5", "-N0", "5", "-s", "1", "-tf", "0.05"]),
    ("02_hyperbolic_02_var_coeff_coarse_grid", ["-NM", "5", "-N0", "5", "-N1", "4", "-N2", "5", "-tf", "0.05"]),
    ("02_hyperbolic_05_burgers_adapt", ["-NM", "5", "-N0", "3", "-tf", "0.02", "-p", "1000"]),
    ("02_hyperbolic_07_kpp", ["-NM", "5", "-N0", "5", "-tf", "0.005", "-p", "1000"]),
]


@pytest.mark.parametrize("driver,args", CASES)
def test_reference_driver_runs_on_gpu(driver, args, tmp_path):
    exe = os.path.join(FACADE, "ref_" + driver)
    if not os.access(exe, os.X_OK):
        pytest.skip("reference driver not built (the reference tree was absent at build time)")
    print(f"[case] {driver} {' '.join(args)}", flush=True)
    r = subprocess.run([exe] + args, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    print(out[-2000:], flush=True)
    assert r.returncode == 0, out[-2000:]
    assert not re.search(r"\bnan\b|\binf\b", out, re.IGNORECASE), "non-finite value printed"


def test_reference_config1_driver_with_lagrange_degree_5_runs():
    # 02_hyperbolic_01_scalar_const_coefficient sets LagrBasis::PMAX = 5
    # (msh_case 1) for its adaptive initial interpolation, then steps the
    # assembled linear operator with RK3
    exe = os.path.join(FACADE, "ref_02_hyperbolic_01_scalar_const_coefficient")
    if not os.access(exe, os.X_OK):
        pytest.skip("reference driver not built (the reference tree was absent at build time)")
    print("[case] 02_hyperbolic_01_scalar_const_coefficient -N 5 -tf 0.05", flush=True)
    r = subprocess.run([exe, "-N", "5", "-tf", "0.05"], capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    print(out[-2000:], flush=True)
    assert r.returncode == 0, out[-2000:]
    assert "evolution finished" in out
    assert not re.search(r"\bnan\b|\binf\b", out, re.IGNORECASE), "non-finite value printed"
