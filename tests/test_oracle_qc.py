"""Pin the oracle's quantum-correction effective potential (src/mcmodel.cpp:2337-2662).

- On a mesh that is silicon throughout, every node is a source and the only
  box edges cut short are at the mesh faces, so the sum separates into one sum
  of erf differences per axis, and each of those telescopes (box i's high edge
  is box i+1's low edge): a constant potential c gives
  c/8 * prod_axes [erf(E_hi - q) - erf(E_lo - q)] with the window's outer edges.
- The window of a node is the reference's iterator walk: one grid line below
  the first line at or above l - r, through the first line above l + r.
- Oxide-only nodes add the barrier -Eb; with no oxide in reach the result does
  not depend on Eb, near oxide it moves linearly with it.
- The correction is linear in the potential once the barrier part is removed,
  and leaves every node outside the quantum region unchanged.
"""
import ctypes as C
import os
import re
import shutil

import numpy as np
import pytest
from scipy.special import erf

import oracle_lib
from helpers import cfg1_deck, emcgpu

SQRT2 = 1.4142135623730951


def qc_deck(tmp_path, name, all_silicon=False, ep="4 0.5 4 0.5 4 0.5 3.1"):
    src = os.path.dirname(cfg1_deck(qc=1, quantum_start_step=2))
    d = tmp_path / name
    shutil.copytree(src, d)
    ldg = (d / "ldg.txt").read_text()
    ldg = re.sub(r"ep_parm [^\n]*", "ep_parm " + ep, ldg)
    if all_silicon:
        ldg = ldg.replace("region -1 19 -20 20 -3 9 VACUUM", "region -1 19 -20 20 -3 9 SILICON")
        ldg = "\n".join(l for l in ldg.split("\n") if not (l.startswith("region") and l.endswith("OXIDE")))
        ldg = ldg.replace("surface_scatter_range 0 10 -20 20 0 5",
                          "surface_scatter_range 0 10 -20 20 0 5\nquantumRegion -1 19 -20 20 -3 9")
    (d / "ldg.txt").write_text(ldg)
    return str(d / "input.txt")


def qc(eng, pot):
    out = np.zeros(eng.mesh.nnode)
    pot = np.ascontiguousarray(pot, np.float64)
    assert oracle_lib.load().orc_qc_potential(C.byref(eng.mesh), oracle_lib.dptr(pot), oracle_lib.dptr(out)) == 0
    return out


def window(l, r):
    """The reference's walk (src/mcmodel.cpp:2584-2625), index by index."""
    lo, hi, a, b = [], [], 0, 0
    for p in range(len(l)):
        while a < p and l[a] < l[p] - r:
            a += 1
        while b < len(l) and not l[b] > l[p] + r:
            b += 1
        if a:
            a -= 1
        if b == len(l):
            b -= 1
        lo.append(a)
        hi.append(b)
    return lo, hi


def test_constant_potential_telescopes(tmp_path):
    e = emcgpu.Engine(qc_deck(tmp_path, "si", all_silicon=True), device=-1)
    try:
        m = e.mesh
        ma = e.mesh_arrays()
        assert ma["qcregion"].all()
        c = 0.37
        out = qc(e, np.full(m.nnode, c))
        ls = [ma["lx"], ma["ly"], ma["lz"]]
        ds = [ma["dx"], ma["dy"], ma["dz"]]
        fac = []
        for a in range(3):
            l, dl, th, ra = ls[a], ds[a], m.qc_theta[a], m.qc_ratio[a]
            r = ra * th
            lo, hi = window(l, r)
            n = len(l)
            f = np.zeros(n)
            for p in range(n):
                e1 = l[lo[p]] - r if lo[p] == 0 else l[lo[p]] - 0.5 * dl[lo[p] - 1]
                e2 = l[hi[p]] + r if hi[p] == n - 1 else l[hi[p]] + 0.5 * dl[hi[p]]
                q = l[p] / (SQRT2 * th)
                f[p] = erf(e2 / (SQRT2 * th) - q) - erf(e1 / (SQRT2 * th) - q)
            fac.append(f)
        want = np.zeros(m.nnode)
        for j in range(m.pny):
            for i in range(m.pnx):
                want[(j * m.pnx + i) * m.pnz:(j * m.pnx + i + 1) * m.pnz] = c / 8 * fac[0][i] * fac[1][j] * fac[2]
        np.testing.assert_allclose(out, want, rtol=1e-13, atol=0)
        # deep inside, the window holds +-r and the Gaussian almost all of its mass
        assert np.all(np.abs(out - c) < 1e-3 * c)
    finally:
        e.close()


def test_barrier_and_linearity(tmp_path):
    a = emcgpu.Engine(qc_deck(tmp_path, "eb31"), device=-1)
    b = emcgpu.Engine(qc_deck(tmp_path, "eb62", ep="4 0.5 4 0.5 4 0.5 6.2"), device=-1)
    try:
        m = a.mesh
        assert b.mesh.qc_eb == pytest.approx(2 * m.qc_eb, rel=1e-15)
        rng = np.random.default_rng(4)
        u, v = rng.normal(size=m.nnode), rng.normal(size=m.nnode)
        z = np.zeros(m.nnode)
        base_a, base_b = qc(a, z), qc(b, z)
        quantum = base_a != 0
        # the barrier part doubles with Eb, and reaches nodes near the oxide only
        np.testing.assert_allclose(base_b, 2 * base_a, rtol=1e-12, atol=1e-12)
        assert 0 < quantum.sum() < np.count_nonzero(qc(a, np.ones(m.nnode)) != 1.0)
        # linear in the potential once the barrier part is removed
        lin = qc(a, 2 * u - 3 * v) - base_a
        np.testing.assert_allclose(lin, 2 * (qc(a, u) - base_a) - 3 * (qc(a, v) - base_a), rtol=1e-9, atol=1e-9)
        # nodes outside the quantum region keep the Poisson potential
        ma = a.mesh_arrays()
        qn = np.zeros(m.nnode, bool)
        cq = np.nonzero(ma["qcregion"])[0]
        k = cq % m.cnz
        t = cq // m.cnz
        i, j = t % m.cnx, t // m.cnx
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    qn[((j + dj) * m.pnx + i + di) * m.pnz + k + dk] = True
        out = qc(a, u)
        np.testing.assert_array_equal(out[~qn], u[~qn])
        assert np.all(out[qn] != u[qn])
    finally:
        a.close()
        b.close()
