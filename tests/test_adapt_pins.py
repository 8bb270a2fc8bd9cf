"""Refine / coarsen of the oracle against an independent set-based statement of
the reference rules (DGAdapt::refine / coarsen, source/DGAdapt.cpp:371-404,
675-688, 982-1016, 1250-1385)."""
import math

import numpy as np
import pytest

import pyoracle


def children(lv, sp, nmax):
    out = []
    for d in range(len(lv)):
        l2 = list(lv)
        l2[d] += 1
        if lv[d] == 0:
            out.append((tuple(l2), tuple(sp)))
        elif lv[d] < nmax:
            for s in (2 * sp[d] - 1, 2 * sp[d] + 1):
                s2 = list(sp)
                s2[d] = s
                out.append((tuple(l2), tuple(s2)))
    return out


def parents(lv, sp):
    out = []
    for d in range(len(lv)):
        if lv[d] == 0:
            continue
        l2, s2 = list(lv), list(sp)
        l2[d] -= 1
        h = (sp[d] + 1) // 2
        s2[d] = h if h % 2 == 1 else h - 1
        out.append((tuple(l2), tuple(s2)))
    return out


def elements(o):
    lev, sup, _ = o.grid()
    return [(tuple(map(int, lev[e])), tuple(map(int, sup[e]))) for e in range(o.num_elem)]


def norms(o):
    u = o.get(pyoracle.FIELD_UCOE).reshape(o.num_elem, -1)
    return [math.pow(sum(math.pow(x, 2.0) for x in row), 0.5) for row in u]


def expected_refine(elems, nrm, eps, nmax):
    S = set(elems)
    new = set(S)
    for el, n in zip(elems, nrm):
        ch = children(*el, nmax)
        if sum(c in S for c in ch) < len(ch) and n > eps:
            new.update(ch)
    work = list(new - S)
    while work:
        el = work.pop()
        for p in parents(*el):
            if p not in new:
                new.add(p)
                work.append(p)
    return new


def expected_coarsen(elems, nrm, eta, nmax):
    alive = dict(zip(elems, nrm))
    while True:
        z = [el for el, n in alive.items()
             if not any(c in alive for c in children(*el, nmax)) and n < eta and max(el[0]) > 0]
        if not z:
            return set(alive)
        for el in z:
            del alive[el]


@pytest.mark.parametrize("dim,nmax,seed", [(2, 6, 0), (2, 7, 1), (3, 4, 2)])
def test_refine_then_coarsen_match_set_rules(dim, nmax, seed):
    rng = np.random.default_rng(seed)
    o = pyoracle.Oracle(dim, 1, 1, nmax, True, 3)
    K = o.K
    for _ in range(3):
        u = rng.standard_normal(o.num_elem * K) * np.repeat(10.0 ** rng.uniform(-6, 0, o.num_elem), K)
        o.set(pyoracle.FIELD_UCOE, u)
        elems, nrm = elements(o), norms(o)
        want = expected_refine(elems, nrm, 1e-2, nmax)
        o.refine(1e-2)
        assert set(elements(o)) == want
        # no holes: every parent of every element is present
        have = set(elements(o))
        assert all(p in have for el in have for p in parents(*el))
        u = rng.standard_normal(o.num_elem * K) * np.repeat(10.0 ** rng.uniform(-6, 0, o.num_elem), K)
        o.set(pyoracle.FIELD_UCOE, u)
        elems, nrm = elements(o), norms(o)
        want = expected_coarsen(elems, nrm, 1e-3, nmax)
        o.coarsen(1e-3)
        assert set(elements(o)) == want
