"""State files on the GPU: runs reproduce the oracle's golden states byte for
byte, reading a state and writing it back is the identity (single solves and
batches), a run continued from a read state matches, and malformed files are
rejected with an error."""
import os
import sys

import numpy as np
import pytest

import pyadg
import workloads
from state_file import unpack

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_state_fixtures as fx  # noqa: E402

pytestmark = pytest.mark.gpu


def golden(name, which):
    with open(fx.path(name, which), "rb") as f:
        return f.read()


def written(sol, tmp_path, tag):
    p = tmp_path / f"{tag}.adgstate"
    sol.state_write(str(p))
    return p.read_bytes()


@pytest.mark.parametrize("name", sorted(fx.CASES))
@pytest.mark.parametrize("form", [0, 1])
def test_gpu_run_writes_the_golden_states(name, form, tmp_path):
    print(f"[case] golden state {name} form={form}", flush=True)
    g = pyadg.Solution(2, 2, 3, 5, True, 3)
    g.init_adaptive_intp_builtin(fx.case_ic(), 1e-4)
    g.set_interpolation(fx.CASES[name]["family"])
    assert written(g, tmp_path, "t0") == golden(name, "t0")
    prob = fx.case_problem()
    prob.form = form
    g.run_hyperbolic(prob)
    assert written(g, tmp_path, "final") == golden(name, "final")


@pytest.mark.parametrize("name", sorted(fx.CASES))
def test_run_continued_from_a_read_state(name, tmp_path):
    print(f"[case] continue from state {name}", flush=True)
    g = pyadg.Solution(2, 2, 3, 5, True, 0)
    g.state_read(fx.path(name, "t0"))
    g.set_interpolation(fx.CASES[name]["family"])
    assert written(g, tmp_path, "t0") == golden(name, "t0")
    g.run_hyperbolic(fx.case_problem())
    assert written(g, tmp_path, "final") == golden(name, "final")


def test_batch_state_round_trip(tmp_path):
    print("[case] batch state round trip", flush=True)
    ws = [workloads.c2_burgers(seed=s, nmax=5, level_init=3, t_final=0.01) for s in (4, 9, 2)]
    b = workloads.make_batch(ws)
    prob = ws[0].problem
    prob.form = 1
    b.run_hyperbolic(prob)
    data = written(b, tmp_path, "batch")
    hdr, insts = unpack(data)
    assert len(insts) == 3
    for i in range(3):
        lv, sp, _ = b.instance_grid(i)
        assert np.array_equal(insts[i][0], lv) and np.array_equal(insts[i][1], sp)
        assert np.array_equal(insts[i][2], b.instance_get(i))
    c = pyadg.Solution(2, 2, 3, 5, True, 0)
    c.state_read(str(tmp_path / "batch.adgstate"))
    assert c.num_instances == 3
    assert written(c, tmp_path, "again") == data


def test_malformed_state_files_are_rejected(tmp_path):
    print("[case] malformed state files", flush=True)
    good = golden("burgers_lagr_2d_n5", "final")
    g = pyadg.Solution(2, 2, 3, 5, True, 3)
    before = written(g, tmp_path, "before")
    cases = {"magic": b"XDGSTAT1" + good[8:], "truncated": good[:-8], "trailing": good + b"\0",
             "level": good[:8 + 24 + 8] + np.array([9, 0, 1, 1], dtype="<i4").tobytes() + good[8 + 24 + 8 + 16:]}
    for tag, data in cases.items():
        p = tmp_path / f"{tag}.adgstate"
        p.write_bytes(data)
        with pytest.raises(pyadg.AdgError, match="state file"):
            g.state_read(str(p))
    other = pyadg.Solution(2, 1, 3, 5, True, 3)
    with pytest.raises(pyadg.AdgError, match="differ"):
        other.state_read(fx.path("burgers_lagr_2d_n5", "final"))
    # a rejected read leaves nothing half-applied when it fails before the grid is replaced
    assert written(g, tmp_path, "after") == before
