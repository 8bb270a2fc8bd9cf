"""Pins of the oracle's level-sum cap (the reference's _coarse_grid variants,
source/FastMultiplyLU.cpp:514-594, source/BilinearForm.cpp:89-159).

On a sparse grid of level N the elements with level sum <= N1 are exactly the
sparse grid of level N1, so every capped operator on grid N must equal the
uncapped operator on grid N1 on those elements and give zero on the others.
CPU only."""
import numpy as np
import pytest

import pyoracle

FIELD_FUCOE = 10


def _grids(dim, k, m, n, n1):
    big = pyoracle.Oracle(dim, k, m, n, True, n)
    small = pyoracle.Oracle(dim, k, m, n, True, n1)
    lb, _, kb = big.grid()
    ls, _, ks = small.grid()
    keep = lb.sum(axis=1) <= n1
    assert np.array_equal(kb[keep], ks), "level-sum subset of grid N is not grid N1"
    return big, small, keep


def _expand(mask, per):
    return np.repeat(mask, per)


@pytest.mark.parametrize("dim,k,m,n,n1", [(2, 2, 3, 5, 3), (3, 1, 2, 4, 2)])
def test_capped_operators_equal_operators_on_the_coarse_sparse_grid(dim, k, m, n, n1):
    rng = np.random.default_rng(7)
    big, small, keep = _grids(dim, k, m, n, n1)
    K, P = (k + 1) ** dim, (m + 1) ** dim
    u = rng.standard_normal(big.num_elem * K)
    big.set(pyoracle.FIELD_UCOE, u)
    small.set(pyoracle.FIELD_UCOE, u[_expand(keep, K)])
    big.set_level_sum_cap(n1)

    big.eval_up(); small.eval_up()
    a, b = big.get(pyoracle.FIELD_UP), small.get(pyoracle.FIELD_UP)
    assert np.array_equal(a[_expand(keep, P)], b)
    assert not np.any(a[_expand(~keep, P)])

    for d in range(dim):
        f = rng.standard_normal(big.num_elem * P)
        big.set(FIELD_FUCOE + d, f)
        small.set(FIELD_FUCOE + d, f[_expand(keep, P)])
    for op in ("rhs_lagr_vol", "rhs_lagr_flx", "penalty", "linear", "linear_assembled"):
        big.rhs_zero(); small.rhs_zero()
        for o in (big, small):
            if op == "penalty":
                o.rhs_penalty([1.3] * dim)
            elif op == "linear":
                o.rhs_linear([1.0, -0.5, 0.75][:dim])
            elif op == "linear_assembled":
                o.rhs_linear([1.0, -0.5, 0.75][:dim], assembled=True)
            else:
                getattr(o, op)()
        a, b = big.get(pyoracle.FIELD_RHS), small.get(pyoracle.FIELD_RHS)
        assert np.array_equal(a[_expand(keep, K)], b), op
        assert not np.any(a[_expand(~keep, K)]), op


def test_cap_at_or_above_the_grid_is_no_cap():
    rng = np.random.default_rng(3)
    a = pyoracle.Oracle(2, 2, 3, 4, True, 4)
    b = pyoracle.Oracle(2, 2, 3, 4, True, 4)
    u = rng.standard_normal(a.num_elem * 9)
    a.set(pyoracle.FIELD_UCOE, u); b.set(pyoracle.FIELD_UCOE, u)
    a.set_level_sum_cap(4)
    a.eval_up(); b.eval_up()
    assert np.array_equal(a.get(pyoracle.FIELD_UP), b.get(pyoracle.FIELD_UP))
