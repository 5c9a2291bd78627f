"""The point-location octree that answers Getisym's list scans (src/ParticleUtils.cpp:
395-625), checked on the CPU against the oracle's literal scans.

The library's Getisym (the code the flight kernels run, here on the host tables through
emc_band_locate) must return the same tetrahedron and symmetry operation as the oracle for
points that exercise every branch of the scan order: the hint tetrahedron, the tlists
bin, the tlist bins ie, ie+1 and ie-1, the scan over all tetrahedra, and no tetrahedron
at all.  Points on shared vertices, edges and faces lie in several tetrahedra at once, so
they pin which one of them each scan returns.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(), device=-1)
    e.tabs = e.band_arrays()
    yield e
    e.close()


def out_of_wedge(b, isym, kw):
    """k of the wedge point kw under symmetry operation isym (OutWedge)."""
    ms = b["matsym"].reshape(6, 48)
    return np.array([kw[ms[r, isym] - 1] * ms[3 + r, isym] for r in range(3)])


def oracle_locate(eng, lib, k, E, band, hint):
    isym, itet = C.c_int(0), C.c_int(0)
    kk = np.ascontiguousarray(k, np.float64)
    st = lib.orc_getisym(C.byref(eng.band), kk.ctypes.data_as(C.POINTER(C.c_double)), float(E), int(band), int(hint),
                         C.byref(isym), C.byref(itet))
    return (-1 if st != 0 else itet.value), isym.value


def corners(b, it):
    nt = len(b["ibt"])
    tet = b["tet"].reshape(4, nt)
    return np.array([[b["xkk"][tet[c, it]], b["ykk"][tet[c, it]], b["zkk"][tet[c, it]]] for c in range(4)])


def points(eng, n, seed):
    """(k, energy, band, hint, kind) samples covering every branch of the scan order."""
    b = eng.tabs
    nt = len(b["ibt"])
    rng = np.random.default_rng(seed)
    dl = 1.0 / eng.band.dlist
    out = []
    for q in range(n):
        it = int(rng.integers(0, nt))
        X = corners(b, it)
        kind = q % 5
        if kind == 0:    # interior point
            w = rng.dirichlet(np.ones(4) * 2)
        elif kind == 1:  # a corner, shared by many tetrahedra
            w = np.eye(4)[rng.integers(0, 4)]
        elif kind == 2:  # an edge midpoint or a face centroid
            m = rng.choice(4, size=int(rng.integers(2, 4)), replace=False)
            w = np.zeros(4)
            w[m] = 1.0 / len(m)
        else:            # interior point, off-bucket energy or a point off the tables
            w = rng.dirichlet(np.ones(4))
        kw = w @ X
        s = int(rng.integers(0, 48))
        band = int(b["ibt"][it])
        e_corner = b["eek"][b["tet"].reshape(4, nt)[:, it]]
        E = float(w @ e_corner)
        if kind == 3:    # energy shifted by up to a few coarse bins: ie+-1 and the full scan
            E += dl * float(rng.choice([-3.0, -1.0, -0.5, 0.5, 1.0, 3.0])) * rng.uniform(0.5, 1.0)
        if kind == 4:    # a random band, or a point beyond the tables
            band = int(rng.integers(0, 8))
            kw = kw + rng.normal(0, 0.05, 3) if q % 2 else kw * 4.0 + 3.0
        hint = int(rng.integers(0, nt)) if rng.random() < 0.2 else 0
        out.append((out_of_wedge(b, s, kw), E, band, hint, kind))
    return out


def scan_branch(eng, t, E, band):
    """Which of Getisym's scans returns tetrahedron t: 0 tlists bin, 1-3 tlist bins
    ie, ie+1, ie-1, 4 the scan over all tetrahedra (t must not be in an earlier one)."""
    b = eng.tabs
    bv = eng.band

    def bucket(n, p, lst, ie):
        return set(lst[p[ie * 8 + band]:p[ie * 8 + band] + n[ie * 8 + band]].tolist())

    ies = max(int(E * bv.dlists), 0)
    ie = min(max(int(E * bv.dlist), 0), 324)
    if ies < 60 and t in bucket(b["ntlists"], b["ptlists"], b["tlists"], ies):
        return 0
    for br, j in ((1, ie), (2, ie + 1), (3, ie - 1)):
        if 0 <= j < 325 and t in bucket(b["ntlist"], b["ptlist"], b["tlist"], j):
            return br
    return 4


def test_index_is_built(eng):
    assert eng.band.tindex_nodes > 0
    assert eng.band.tindex_entries >= eng.band.nt


@pytest.mark.parametrize("seed", [1, 2])
def test_locate_equals_oracle_scans(eng, seed):
    lib = oracle_lib.load()
    pts = points(eng, 1500, seed)
    k = np.array([p[0] for p in pts])
    E = np.array([p[1] for p in pts])
    band = np.array([p[2] for p in pts], np.int32)
    hint = np.array([p[3] for p in pts], np.int32)
    got_t, got_s = eng.band_locate(k, E, band, hint)
    kinds = np.array([p[4] for p in pts])
    found = {kd: 0 for kd in range(5)}
    branches = np.zeros(5, int)
    for q in range(len(pts)):
        want_t, want_s = oracle_locate(eng, lib, k[q], E[q], band[q], hint[q])
        assert (got_t[q], got_s[q]) == (want_t, want_s), f"point {q} kind {kinds[q]}"
        found[kinds[q]] += want_t >= 0
        if want_t >= 0 and want_t != hint[q]:
            branches[scan_branch(eng, want_t, E[q], band[q])] += 1
    # every kind of point was located at least sometimes, and kind 4 also missed
    assert all(v > 0 for v in found.values())
    assert found[4] < (kinds == 4).sum()
    # every scan of the search order produced some of the answers
    assert np.all(branches > 0), branches
