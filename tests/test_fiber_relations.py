"""The device kernels enumerate related fiber cells by index arithmetic
(csrc/fiber.hpp).  The reference builds the same sets from interval
predicates (Element::is_vol_alpt / is_flx_alpt, source/Element.cpp:265-348);
the oracle keeps those predicate-built sets.  Here the two must agree, element
by element, in the same order, on sparse, full and adaptively refined grids."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle
from common import ic_params

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def enum_lib():
    d = tempfile.mkdtemp()
    so = os.path.join(d, "fiber_enum.so")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-o", so,
                           os.path.join(HERE, "helpers", "fiber_enum.cpp")])
    L = ctypes.CDLL(so)
    L.fiber_related.restype = ctypes.c_int
    L.fiber_related.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
    L.fiber_present.restype = ctypes.c_int
    L.fiber_present.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int,
                                                     ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
    return L


def slot(level, suppt):
    return 0 if level == 0 else (1 << (level - 1)) + (suppt - 1) // 2


def check_grid(o, L):
    lev, sup, _ = o.grid()
    E, D = lev.shape
    slots = np.array([[slot(lev[e, d], sup[e, d]) for d in range(D)] for e in range(E)])
    buf = np.zeros(1 << 16, np.int32)
    for d in range(D):
        others = [dd for dd in range(D) if dd != d]
        fibers = {}
        for e in range(E):
            fibers.setdefault(tuple(slots[e, others]), {})[slots[e, d]] = e
        for e in range(E):
            fib = fibers[tuple(slots[e, others])]
            maxl = max(int(lev[x, d]) for x in fib.values())
            for flx in (0, 1):
                ref = list(o.neighbors(e, d, flx))
                for lu in (0, 1, 2):
                    n = L.fiber_related(int(slots[e, d]), maxl, flx, lu,
                                        buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), buf.size)
                    got = [fib[s] for s in buf[:n] if s in fib]
                    lt = lev[e, d]
                    if lu == 1:
                        want = [x for x in ref if lev[x, d] > lt]
                    elif lu == 2:
                        want = [x for x in ref if lev[x, d] <= lt]
                    else:
                        want = ref
                    assert got == want, (e, d, flx, lu)
                    assert len(set(buf[:n])) == n  # no slot enumerated twice


@pytest.mark.parametrize("dim,nmax,sparse", [(1, 6, False), (2, 5, True), (2, 4, False), (3, 4, True)])
def test_relations_regular_grids(enum_lib, dim, nmax, sparse):
    o = pyoracle.Oracle(dim, 1, 1, nmax, sparse, nmax)
    check_grid(o, enum_lib)


def test_relations_adaptive_grid(enum_lib):
    o = pyoracle.Oracle(2, 2, 3, 7, True, 3)
    p = ic_params(2, base=0.5, ampl=0.25, bumps=[(0.2, 0.05, (0.1, 0.93)), (0.1, 0.08, (0.8, 0.2))])
    o.init_adaptive_intp(lambda x: pyadg_ic(p, x), 1e-5)
    assert o.num_elem > 300
    check_grid(o, enum_lib)


def pyadg_ic(p, x):
    s = x.sum()
    v = p.base + p.ampl * np.sin(2 * np.pi * s)
    for b in range(p.nbump):
        c = np.array([p.bump_center[b][d] for d in range(len(x))])
        t = np.abs(x - c)
        t = np.minimum(t, 1 - t)
        v += p.bump_amp[b] * np.exp(-(t @ t) / p.bump_width[b] ** 2)
    return v


@pytest.mark.parametrize("nmax", [3, 5, 8])
def test_present_enumeration_matches_filtered_relations(enum_lib, nmax):
    # the occupancy-bit walk yields exactly the present related slots, in ascending order,
    # for fibers that start anywhere inside a word of bits
    rng = np.random.default_rng(nmax)
    ns = 1 << nmax
    f_rel = enum_lib.fiber_related
    f_pre = enum_lib.fiber_present
    for trial in range(40):
        fb = int(rng.integers(0, 96))
        present = rng.random(ns) < rng.choice([0.1, 0.5, 0.9])
        bits = np.zeros((fb + ns + 31) // 32 + 1, np.uint32)
        for s in np.nonzero(present)[0]:
            b = fb + int(s)
            bits[b >> 5] |= np.uint32(1 << (b & 31))
        bits_p = bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        out = np.zeros(4 * ns, np.int32)
        out_p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        for st in rng.integers(0, ns, 12):
            for flx in (0, 1):
                for lu in (0, 1, 2):
                    maxl = int(rng.integers(0, nmax + 1))
                    n = f_rel(int(st), maxl, flx, lu, out_p, out.size)
                    want = [int(s) for s in out[:n] if present[s]]
                    n = f_pre(int(st), maxl, flx, lu, bits_p, fb, out_p, out.size)
                    got = [int(s) for s in out[:n] if present[s]]
                    assert got == want, (st, flx, lu, maxl)
