"""Pins of the oracle's artificial viscosity in its matrix-free form
(DGAdapt::update_viscosity_element, source/DGAdapt.cpp:727-759;
LargViscosInterpolation::support_gradu_Lagr_fast, source/Interplation.cpp:
4176-4189,579-640): the shock detector against hand-set norms, and the
interpolated viscous flux against the closed-form derivative of a cubic.  CPU only."""
import numpy as np
import pytest

import pyoracle

G = np.polynomial.Polynomial([0.3, -1.0, 0.8, 0.5])
H = np.polynomial.Polynomial([1.0, 0.4, -0.6, 0.2])


def _order(level, suppt, p, pmax):
    slot = 0 if level == 0 else (1 << (level - 1)) + (suppt - 1) // 2
    return slot * (pmax + 1) + p


def test_shock_detector_by_hand():
    # full grid of max level 3 in 2D: the childless elements at the finest level are the (3, 3) ones
    k = 2
    o = pyoracle.Oracle(2, k, 3, 3, False, 3)
    lev, sup, _ = o.grid()
    K = (k + 1) ** 2
    u = np.zeros(o.num_elem * K)
    fine = [e for e in range(o.num_elem) if tuple(lev[e]) == (3, 3)]
    assert len(fine) == 16
    bound = 2.0 ** ((1 - k) * 6)
    # kappa 0.5: selected iff log10(norm / bound) > 0.5
    u[fine[0] * K] = bound * 10 ** 0.6
    u[fine[1] * K + 3] = bound * 10 ** 0.4
    u[fine[2] * K:(fine[2] + 1) * K] = bound * 10 ** 0.55 / 3.0  # norm = bound * 10^0.55
    coarse = [e for e in range(o.num_elem) if tuple(lev[e]) == (3, 2)][0]
    u[coarse * K] = 1.0  # large, but it has children: never selected
    o.set(pyoracle.FIELD_UCOE, u)
    n, flags = o.viscosity_elements(0.5)
    assert n == 2 and flags[fine[0]] == 1 and flags[fine[2]] == 1 and flags.sum() == 2


@pytest.mark.parametrize("overlap", [False, True])
def test_viscous_flux_is_the_derivative_on_the_viscosity_support(overlap):
    o = pyoracle.Oracle(2, 3, 3, 3, False, 3)
    o.init_separable([lambda x, d: float(G(x)) if d == 0 else float(H(x))], 1e10)
    n, flags = o.viscosity_elements(-100.0)  # every childless finest element
    assert n == 16
    coef = -0.37
    o.viscosity_flux(coef, overlap)
    lev, sup, _ = o.grid()
    fine = [(lev[e], sup[e]) for e in range(o.num_elem) if flags[e]]
    for d in range(2):
        fp = o.get(pyoracle.FIELD_FP + d).reshape(-1, 16)
        for e in range(o.num_elem):
            for q in range(16):
                pq = (q // 4, q % 4)
                x = [o.lagr_point(_order(lev[e, dd], sup[e, dd], pq[dd], 3)) for dd in range(2)]
                du = G.deriv()(x[0]) * H(x[1]) if d == 0 else G(x[0]) * H.deriv()(x[1])
                if overlap:
                    # number of level-(3, 3) boxes (width 1/4, closed) that hold the point
                    cnt = 1
                    for dd in range(2):
                        c = 4 * x[dd]
                        cnt *= 2 if abs(c - round(c)) < 1e-15 and 0 < round(c) < 4 else 1
                    want = cnt * coef * du
                else:
                    want = coef * du
                assert abs(fp[e, q] - want) < 1e-11 * max(1, abs(want)), (d, lev[e], sup[e], q, fp[e, q], want)
