"""Pin the oracle's deposit, RHS, field, Rho and Poisson restatements by closed
forms and invariants (charge conservation, corner placement, exact fields of a
linear potential, manufactured Poisson solutions, operator symmetry)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu

FXQ = emcgpu.FX_Q


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(), device=-1)
    e.particles_init()
    e.pop0 = e.download()
    e.marr = e.mesh_arrays()
    return e


def test_cic_conserves_charge(eng):
    p = eng.pop0
    node_q, node_stat, cell_q, cell_n = oracle_lib.deposit(eng, p, p.n)
    nc = eng.mesh.ncell
    # each of the 8 rounded corner shares is off by at most half a unit
    total = p.charge[:p.n].sum() * FXQ
    assert abs(node_q.sum() - total) <= 8 * p.n
    assert cell_n.sum() == p.n
    assert abs(cell_q[:nc].sum() - p.charge[p.ptype == 0].sum() * FXQ) <= p.n
    assert abs(cell_q[nc:2 * nc].sum() - p.charge[p.ptype == 1].sum() * FXQ) <= p.n
    assert abs(cell_q[2 * nc:].sum() - np.abs(p.charge).sum() * FXQ) <= p.n
    # |q| statistic equals the |q| sum of the CIC weights (statistic, src/mcmodel.cpp:1167-1408)
    nn = eng.mesh.nnode
    qabs_e = node_stat[3 * nn:4 * nn].sum() / emcgpu.FX_STAT
    assert abs(qabs_e - np.abs(p.charge[p.ptype == 0]).sum()) < 1e-6 * max(1, qabs_e)


def test_particle_on_a_node_deposits_there(eng):
    p = eng.pop0.subset(np.arange(3))
    m = eng.marr
    p.x[:] = m["lx"][p.i]
    p.y[:] = m["ly"][p.j]
    p.z[:] = m["lz"][p.k]
    node_q, _, _, _ = oracle_lib.deposit(eng, p, 3)
    expect = np.zeros_like(node_q)
    for q in range(3):
        n = (p.j[q] * eng.mesh.pnx + p.i[q]) * eng.mesh.pnz + p.k[q]
        expect[n] += np.int64(np.rint(p.charge[q] * FXQ))
    # each particle's whole charge lands on its own node and nowhere else
    assert np.array_equal(node_q, expect)
    nz = np.nonzero(node_q)[0]
    assert len(nz) <= 3
    assert node_q.sum() == np.rint(p.charge[:3] * FXQ).sum()


def test_field_of_linear_potential_is_exact(eng):
    """pot2field (src/mcmodel.cpp:2776-2813) reproduces the constant field of a
    linear potential."""
    lib = oracle_lib.load()
    m = eng.mesh
    a = eng.marr
    X, Y, Z = np.meshgrid(a["lx"], a["ly"], a["lz"], indexing="ij")
    # node order (j, i, k)
    pot = (0.3 * X - 1.7 * Y + 0.9 * Z).transpose(1, 0, 2).reshape(-1).copy()
    ex, ey, ez = (np.zeros(m.ncell) for _ in range(3))
    lib.orc_field(C.byref(m), oracle_lib.dptr(pot), oracle_lib.dptr(ex), oracle_lib.dptr(ey), oracle_lib.dptr(ez))
    assert np.allclose(ex, -0.3, rtol=1e-10) and np.allclose(ey, 1.7, rtol=1e-10) and np.allclose(ez, -0.9, rtol=1e-10)


def test_rho_clamp_and_rhs(eng):
    lib = oracle_lib.load()
    m = eng.mesh
    p = eng.pop0
    node_q, _, cell_q, _ = oracle_lib.deposit(eng, p, p.n)
    rho = np.zeros(m.ncell)
    lib.orc_cell_rho(C.byref(m), oracle_lib.i64ptr(cell_q), oracle_lib.dptr(rho))
    da = eng.marr["da"]
    assert np.all(rho >= 0.5 * da)
    dens, rref, relim = (np.zeros(m.nnode) for _ in range(3))
    lib.orc_rhs(C.byref(m), oracle_lib.i64ptr(node_q), oracle_lib.dptr(dens), oracle_lib.dptr(rref),
                oracle_lib.dptr(relim))
    kind = eng.marr["node_kind"]
    np.testing.assert_array_equal(rref[kind == 1], eng.marr["dirichlet"][kind == 1])
    # compute_rhs only fills nodes that carry particle charge (src/mcmodel.cpp:2872-2879)
    assert np.all(rref[(kind == 0) & (dens == 0)] == 0)


def apply_elim(eng, x):
    m, a = eng.mesh, eng.marr
    X = x.reshape(m.pny, m.pnx, m.pnz)
    D, CX, CY, CZ = (a[k].reshape(m.pny, m.pnx, m.pnz) for k in ("diag", "cx", "cy", "cz"))
    y = D * X
    y[:, :-1, :] += CX[:, :-1, :] * X[:, 1:, :]
    y[:, 1:, :] += CX[:, :-1, :] * X[:, :-1, :]
    y[:-1] += CY[:-1] * X[1:]
    y[1:] += CY[:-1] * X[:-1]
    y[:, :, :-1] += CZ[:, :, :-1] * X[:, :, 1:]
    y[:, :, 1:] += CZ[:, :, :-1] * X[:, :, :-1]
    return y.reshape(-1)


def test_elimination_is_symmetric_and_consistent(eng):
    """The eliminated operator is symmetric positive definite and agrees with the
    operator as assembled (init_matrix, src/mcmodel.cpp:4695-4869) on every row."""
    lib = oracle_lib.load()
    m, a = eng.mesh, eng.marr
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=m.nnode), rng.normal(size=m.nnode)
    kind = a["node_kind"]
    x[kind != 0] = 0
    y[kind != 0] = 0
    assert abs(x @ apply_elim(eng, y) - y @ apply_elim(eng, x)) < 1e-10 * abs(x @ apply_elim(eng, y))
    assert x @ apply_elim(eng, x) > 0
    # assembled operator: rows with contact neighbours differ by exactly the moved terms
    phi = rng.normal(size=m.nnode)
    phi[kind == 1] = a["dirichlet"][kind == 1]
    full = np.zeros(m.nnode)
    lib.orc_apply_full(C.byref(m), oracle_lib.dptr(phi), oracle_lib.dptr(full))
    x = phi.copy()
    x[kind != 0] = 0
    el = apply_elim(eng, x)
    u = kind == 0
    np.testing.assert_allclose(el[u] + a["rhs_corr"][u], full[u], rtol=1e-9, atol=1e-9 * np.abs(full).max())
    # stencil rows sum to zero away from contacts and identity rows (flux form)
    st = a["stencil"].reshape(-1, 7)
    interior = u & np.all(st[:, :6] != 0, axis=1)
    assert np.allclose(st[interior].sum(axis=1), 0, atol=1e-12 * np.abs(st).max())


def test_pcg_manufactured_solution(eng):
    lib = oracle_lib.load()
    m, a = eng.mesh, eng.marr
    rng = np.random.default_rng(1)
    kind = a["node_kind"]
    phi = rng.normal(size=m.nnode)
    phi[kind == 1] = a["dirichlet"][kind == 1]
    phi[kind == 2] = 0.0
    full = np.zeros(m.nnode)
    lib.orc_apply_full(C.byref(m), oracle_lib.dptr(phi), oracle_lib.dptr(full))
    b = full.copy()
    b[kind == 0] -= a["rhs_corr"][kind == 0]
    x = np.zeros(m.nnode)
    rr = C.c_double(0)
    it = lib.orc_pcg(C.byref(m), oracle_lib.dptr(b), oracle_lib.dptr(x), 1e-13, 100000, C.byref(rr))
    assert rr.value <= 1e-13 and it > 10
    assert np.max(np.abs(x - phi)) < 1e-7 * np.max(np.abs(phi))
