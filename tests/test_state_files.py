"""Golden state files (tests/golden/*.adgstate): the oracle still writes them
byte for byte, and their layout is the one adg_state_write documents.  CPU only."""
import os
import sys

import numpy as np
import pytest

import pyoracle
from state_file import pack, unpack

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_state_fixtures as fx  # noqa: E402


@pytest.mark.parametrize("name", sorted(fx.CASES))
def test_oracle_reproduces_the_golden_states(name):
    first, final = fx.run_case(name)
    for which, data in (("t0", first), ("final", final)):
        with open(fx.path(name, which), "rb") as f:
            assert f.read() == data, f"{name} {which}: the oracle no longer writes the committed state"


@pytest.mark.parametrize("name", sorted(fx.CASES))
def test_golden_states_are_valid_grids(name):
    for which in ("t0", "final"):
        with open(fx.path(name, which), "rb") as f:
            (dim, k, m, nmax), insts = unpack(f.read())
        assert (dim, k, m, nmax) == (2, 2, 3, 5) and len(insts) == 1
        lv, sp, u = insts[0]
        assert u.size == len(lv) * (k + 1) ** dim and np.all(np.isfinite(u))
        keys = [pyoracle.hash_key(l, s) for l, s in zip(lv, sp)]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
        present = {(tuple(l), tuple(s)) for l, s in zip(lv.tolist(), sp.tolist())}
        for l, s in zip(lv.tolist(), sp.tolist()):
            for d in range(dim):
                if l[d] == 0:
                    continue
                pl, ps = list(l), list(s)
                pl[d] -= 1
                ps[d] = 1 if pl[d] == 0 else (s[d] + 1) // 2 - ((s[d] + 1) // 2 + 1) % 2
                assert (tuple(pl), tuple(ps)) in present, ("hole", l, s)


def test_pack_unpack_round_trip():
    rng = np.random.default_rng(0)
    insts = [(np.array([[0, 0], [1, 0]]), np.array([[1, 1], [1, 1]]), rng.standard_normal(2 * 9)),
             (np.array([[0, 0]]), np.array([[1, 1]]), rng.standard_normal(9))]
    data = pack(2, 2, 3, 5, insts)
    hdr, back = unpack(data)
    assert hdr == (2, 2, 3, 5)
    for a, b in zip(insts, back):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), y)
    assert len(data) == 8 + 24 + 16 + 3 * 2 * 2 * 4 + 3 * 9 * 8
