"""Invariants of the fiber order and the tiles the fiber-tile kernels use,
restated on the host for adaptive grids of the oracle (CPU).

For every dimension d the elements sharing their cells outside d form a
fiber; fibers get consecutive ids in the order of their packed slot keys,
fiber f owns the bits [f * 2^NMAX, (f + 1) * 2^NMAX) of the occupancy map,
and an element's position in fiber order is the rank of its bit.  Tile t
holds the fibers whose first position lies in [t * te, (t + 1) * te).  The
kernels rely on: positions form a permutation of the elements; tiles hold
whole fibers, are disjoint and cover every element once; a tile has at most
te - 1 + (largest fiber) elements; every related source cell present in a
target's fiber (vol and flx relations, every L/U mode, the ascending order the
sums use) has its position inside the target's tile.
"""
import numpy as np
import pytest

import pyoracle
from common import ic_params, init_both  # noqa: F401

LU = (0, 1, 2)


def slot_level(s):
    return 0 if s == 0 else int(s).bit_length()


def related(st, maxl, flx, lu, ns):
    """Related source slots of a target slot, ascending (the enumeration of
    the device kernels, written from the interval predicates)."""
    lt = slot_level(st)
    lo_l = lt + 1 if lu == 1 else 0
    hi_l = min(lt, maxl) if lu == 2 else maxl
    out = []

    def support(s):
        level = slot_level(s)
        if level <= 1:
            return 0.0, 1.0
        n = 1 << (level - 1)
        i = s - n
        return i / n, (i + 1) / n

    tl, tr = support(st)
    for ls in range(lo_l, hi_l + 1):
        cells = [ls] if ls <= 1 else range(1 << (ls - 1), 1 << ls)
        for s in cells:
            sl, sr = support(s)
            inter = min(tr, sr) - max(tl, sl) > 0
            touch = inter or abs(sl - tr) < 1e-13 or abs(sr - tl) < 1e-13 or \
                (abs(sl) < 1e-13 and abs(tr - 1) < 1e-13) or (abs(sr - 1) < 1e-13 and abs(tl) < 1e-13)
            if (touch if flx else inter):
                out.append(s)
    return sorted(out)


def fiber_structure(slots, nmax, d):
    D = slots.shape[1]
    ns = 1 << nmax
    key = [tuple(int(slots[e, q]) if q != d else 0 for q in range(D)) for e in range(len(slots))]
    fids = {k: i for i, k in enumerate(sorted(set(key)))}
    fbase = np.array([fids[k] * ns for k in key], np.int64)
    nbits = len(fids) * ns
    bits = np.zeros(nbits + 32, np.int64)
    bits[fbase + slots[:, d]] = 1
    rank = np.concatenate([[0], np.cumsum(bits)])  # rank[x] = set bits below x
    return fbase, bits, rank, len(fids)


@pytest.mark.parametrize("case", ["burgers2d", "bumps3d"])
@pytest.mark.parametrize("te", [64, 128, 256])
def test_fiber_tiles(case, te):
    if case == "burgers2d":
        dim, k, m, nmax, li = 2, 2, 3, 7, 3
        p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.2, 0.06, (0.3, 0.7)), (0.1, 0.08, (0.8, 0.2))])
    else:
        dim, k, m, nmax, li = 3, 1, 2, 5, 2
        p = ic_params(4, bumps=[(1.0, 0.1, (0.4, 0.5, 0.6))])
    o = pyoracle.Oracle(dim, k, m, nmax, True, li)
    import ctypes
    import pyadg  # noqa: F401  (the library's built-in point function is host code)
    o._keep.append(p)
    o.init_adaptive_intp_c(pyadg.ic_eval_ptr(), ctypes.cast(ctypes.pointer(p), ctypes.c_void_p), 1e-4)
    lv, sp, _ = o.grid()
    slots = np.where(lv == 0, 0, (1 << np.maximum(lv - 1, 0)) + (sp - 1) // 2)
    E = len(slots)
    assert E > 100
    ns = 1 << nmax
    for d in range(dim):
        fbase, bits, rank, nfib = fiber_structure(slots, nmax, d)
        pos = rank[fbase + slots[:, d]]
        assert sorted(pos) == list(range(E)), "positions are not a permutation"
        perm = np.empty(E, np.int64)
        perm[pos] = np.arange(E)
        # fibers contiguous in fiber order, ascending slot inside
        sizes = {}
        for e in range(E):
            sizes[fbase[e]] = sizes.get(fbase[e], 0) + 1
        maxfib = max(sizes.values())
        ntile = (E + te - 1) // te
        tbeg = np.full(ntile, np.iinfo(np.int32).max)
        tend = np.full(ntile, -1)
        for e in range(E):
            first = rank[fbase[e]]
            if pos[e] == first:
                t = pos[e] // te
                tbeg[t] = min(tbeg[t], pos[e])
                tend[t] = max(tend[t], rank[fbase[e] + ns])
        covered = np.zeros(E, np.int64)
        max_tile = te - 1 + maxfib
        for t in range(ntile):
            if tend[t] <= tbeg[t]:
                continue
            assert tend[t] - tbeg[t] <= max_tile
            covered[tbeg[t]:tend[t]] += 1
            for q in range(tbeg[t], tend[t]):
                e = perm[q]
                assert rank[fbase[e]] >= tbeg[t] and rank[fbase[e] + ns] <= tend[t], "fiber split by a tile"
        assert np.all(covered == 1), "tiles do not cover the elements exactly once"
        # related sources of a target lie in its own tile (checked on a sample of targets)
        fmaxl = {}
        for e in range(E):
            fmaxl[fbase[e]] = max(fmaxl.get(fbase[e], 0), slot_level(slots[e, d]))
        rng = np.random.default_rng(d)
        for e in rng.choice(E, size=min(E, 60), replace=False):
            st = int(slots[e, d])
            t = next(t for t in range(ntile) if tbeg[t] <= pos[e] < tend[t])
            for flx in (0, 1):
                for lu in LU:
                    for s in related(st, fmaxl[fbase[e]], flx, lu, ns):
                        x = fbase[e] + s
                        if bits[x]:
                            assert tbeg[t] <= rank[x] < tend[t]
