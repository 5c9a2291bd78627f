"""The oracle's MultipleRefresh (src/MultipleRefresh.cpp:9-143), pinned on the CPU.

gen_new_par_mr keeps a cell's particles of one carrier type when their count is
within a factor Rpar = 2 of the target and sum q^2 < Rsqr (sum q)^2 / n (Rsqr = 2);
otherwise it draws the target number of copies with probability |q| / sum|q| and
gives each the charge sum q / target.  Particles with |q| < pckill = 1e-30 are
dropped (src/mcmodel.cpp:3883-3885).  The tests below fix the keep test at its
boundaries, the draw weights (a binomial test), charge bookkeeping, the output
order and the independence of the input order.
"""
import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(), device=-1)
    e.particles_init()
    e.pop0 = e.download()
    e.marr = e.mesh_arrays()
    yield e
    e.close()


def cell_of(eng, p):
    m = eng.mesh
    return (p.j[:p.n].astype(np.int64) * m.cnx + p.i[:p.n]) * m.cnz + p.k[:p.n]


def synthetic(eng, cell, charges, ptype=0, seed=0):
    """A population of len(charges) particles in one cell, copied from population slots."""
    p0 = eng.pop0
    c0 = cell_of(eng, p0)
    src = np.nonzero((c0 == cell) & (p0.ptype[:p0.n] == ptype))[0]
    assert src.size, "no particle of that type in the cell"
    idx = src[np.arange(len(charges)) % src.size]
    p = p0.subset(idx)
    p.charge[:] = charges
    p.key[:] = np.random.default_rng(seed).integers(1, 2 ** 63, len(charges), dtype=np.uint64)
    return p


def a_silicon_cell(eng):
    c0 = cell_of(eng, eng.pop0)
    return int(np.bincount(c0).argmax())


def targets(eng, cell, ne, nh=1):
    de = np.array(eng.marr["mr_desired_e"], np.int32)
    dh = np.array(eng.marr["mr_desired_h"], np.int32)
    de[cell], dh[cell] = ne, nh
    return de, dh


def test_initial_population_is_kept_in_cell_and_key_order(eng):
    p = eng.pop0
    out = oracle_lib.multiple_refresh(eng, p, p.n, 7)
    assert out.n == p.n
    order = np.lexsort((p.key[:p.n], p.ptype[:p.n], cell_of(eng, p)))
    want = p.subset(order)
    for f in emcgpu.PARTICLE_F64 + emcgpu.PARTICLE_I32 + ("key",):
        np.testing.assert_array_equal(getattr(out, f)[:out.n], getattr(want, f)[:want.n], err_msg=f)


def test_keep_test_boundaries(eng):
    c = a_silicon_cell(eng)
    # equal charges: kept for 2 < n/target... strictly inside (target/2, 2 target)
    for n, target, kept in ((4, 4, True), (3, 4, True), (2, 4, False), (7, 4, True), (8, 4, False)):
        p = synthetic(eng, c, np.full(n, 0.5))
        de, dh = targets(eng, c, target)
        out = oracle_lib.multiple_refresh(eng, p, p.n, 3, de, dh)
        assert out.n == (n if kept else target), (n, target)
        assert kept == bool(np.all(np.isin(out.key[:out.n], p.key))), (n, target)
    # uneven charges fail sum q^2 < 2 (sum q)^2 / n: 1,1,1,100 -> 10003 > 5304.5
    p = synthetic(eng, c, np.array([1.0, 1.0, 1.0, 100.0]))
    de, dh = targets(eng, c, 4)
    out = oracle_lib.multiple_refresh(eng, p, p.n, 3, de, dh)
    assert out.n == 4 and not np.any(np.isin(out.key[:4], p.key))
    np.testing.assert_array_equal(out.charge[:4], np.full(4, 103.0 / 4))


def test_killed_particles_are_dropped(eng):
    c = a_silicon_cell(eng)
    p = synthetic(eng, c, np.array([0.5, 1e-31, 0.5, 0.5, -1e-31]))
    de, dh = targets(eng, c, 3)
    out = oracle_lib.multiple_refresh(eng, p, p.n, 5, de, dh)
    assert out.n == 3
    assert np.all(out.charge[:3] == 0.5)


def test_draw_weights_follow_charge(eng):
    """Two parents with |q| = 1 and 3 and a target of 40 000: copies of the second
    follow Binomial(40000, 3/4)."""
    c = a_silicon_cell(eng)
    p = synthetic(eng, c, np.array([1.0, 3.0]))
    p.x[1] = p.x[0] + 1e-3  # tell the parents apart
    de, dh = targets(eng, c, 40000)
    out = oracle_lib.multiple_refresh(eng, p, p.n, 11, de, dh, cap=40016)
    assert out.n == 40000
    k = np.count_nonzero(out.x[:out.n] == p.x[1])
    assert np.count_nonzero(out.x[:out.n] == p.x[0]) + k == 40000
    sd = np.sqrt(40000 * 0.75 * 0.25)
    assert abs(k - 30000) < 5 * sd, k
    np.testing.assert_array_equal(out.charge[:out.n], np.full(out.n, 4.0 / 40000))
    assert len(np.unique(out.key[:out.n])) == out.n


def test_refresh_bookkeeping_and_input_order(eng):
    """One particle of fifty times the charge in every third cell: those groups are
    resampled, hit their target, carry the group's total charge and copy a parent of
    their own cell and type; the result does not depend on the order of the input
    slots."""
    p = eng.pop0.copy()
    rng = np.random.default_rng(3)
    c = cell_of(eng, p)
    g = 2 * c + p.ptype[:p.n]
    first = np.unique(g, return_index=True)[1]
    pick = first[c[first] % 3 == 0]
    p.charge[pick] *= 50.0
    out = oracle_lib.multiple_refresh(eng, p, p.n, 21)
    perm = rng.permutation(p.n)
    out2 = oracle_lib.multiple_refresh(eng, p.subset(perm), p.n, 21)
    assert out.n == out2.n
    for f in emcgpu.PARTICLE_F64 + emcgpu.PARTICLE_I32 + ("key",):
        np.testing.assert_array_equal(getattr(out, f)[:out.n], getattr(out2, f)[:out2.n], err_msg=f)
    de, dh = eng.marr["mr_desired_e"], eng.marr["mr_desired_h"]
    gi = 2 * c + p.ptype[:p.n]
    go = 2 * cell_of(eng, out) + out.ptype[:out.n]
    tin = np.bincount(gi, weights=p.charge[:p.n], minlength=2 * eng.mesh.ncell)
    tout = np.bincount(go, weights=out.charge[:out.n], minlength=2 * eng.mesh.ncell)
    assert np.all(np.abs(tin - tout) <= 1e-12 * np.maximum(np.abs(tin), 1e-300))
    assert np.all(np.diff(go) >= 0), "output not in cell/type order"
    new = ~np.isin(out.key[:out.n], p.key[:p.n])
    assert new.any()
    ng = np.bincount(go[new], minlength=2 * eng.mesh.ncell)
    tgt = np.where(np.arange(2 * eng.mesh.ncell) % 2 == 0, np.repeat(de, 2), np.repeat(dh, 2))
    assert np.all((ng == 0) | (ng == tgt))
    # every new particle is a copy of a parent of its group (all fields but charge and key)
    fields = ("x", "y", "z", "kx", "ky", "kz", "energy", "itet", "isym")
    have = {(int(g),) + tuple(getattr(p, f)[s] for f in fields) for s, g in enumerate(gi)}
    for s in np.nonzero(new)[0][:2000]:
        assert (int(go[s]),) + tuple(getattr(out, f)[s] for f in fields) in have
