"""GPU, the production FinFET at full size: 80 x 160 x 48 cells (the 20 x 40 x 12
deck refined 4 times, 639 009 nodes), ElectronNumber 2 000 000 and HoleNumber
500 000, 3.13 M sampled particles: the workload bench.py times.  At this size the
oracle is too slow to replay, so the test checks what holds at any size:

- the sampled population lies in the mesh, every particle in a silicon cell of
  its own cell index, with a valid tetrahedron of its carrier type's bands;
- the deposit: each cell's particle count and charge sum equal the particles'
  (int64, exactly: the cell sums are sums of the per-particle fixed-point
  charges), and the node charge equals the particles' charge to the rounding of
  the 8 corner terms of each particle;
- the Poisson solve: the solution of the deposited charge meets the residual
  bound on the operator as the reference assembles it (contact rows included);
- 20 full steps: each step conserves particles and charge through the contact
  tallies (live after = before + generated - caught, charge moved by the
  generated minus the caught charge), drops nothing, and keeps the potential
  finite."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import emcgpu, deck

pytestmark = pytest.mark.gpu

FX_Q = 2.0 ** 48  # fixed-point scale of the deposited charge


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(deck("cfg2", refine=4, electrons=2000000, holes=500000, vd=0.5, vg=0.7), device=0, seed=1)
    yield e
    e.close()


def test_cfg2_population_deposit_and_solve(eng):
    m = eng.mesh
    assert (m.cnx, m.cny, m.cnz, m.nnode) == (80, 160, 48, 639009)
    print("[gpu] cfg2: sampling 3.13 M particles", flush=True)
    n = eng.particles_init()
    assert 3.0e6 < n < 3.3e6
    p = eng.download()
    live = p.i[:p.n] >= 0
    assert int(live.sum()) == n
    i, j, k = p.i[:p.n][live], p.j[:p.n][live], p.k[:p.n][live]
    assert i.min() >= 0 and i.max() < m.cnx and j.min() >= 0 and j.max() < m.cny and k.min() >= 0 and k.max() < m.cnz
    ma = eng.mesh_arrays()
    x, y, z = p.x[:p.n][live], p.y[:p.n][live], p.z[:p.n][live]
    lx, ly, lz = ma["lx"], ma["ly"], ma["lz"]
    assert np.all((x >= lx[i]) & (x <= lx[i + 1]) & (y >= ly[j]) & (y <= ly[j + 1]) & (z >= lz[k]) & (z <= lz[k + 1]))
    SILICON = 3  # material code of the deck grammar's SILICON regions
    assert np.all(ma["material"][(j.astype(np.int64) * m.cnx + i) * m.cnz + k] == SILICON)
    b = eng.band_arrays()
    v = eng.band
    typ = p.ptype[:p.n][live]
    band = b["ibt"][p.itet[:p.n][live]]
    first = np.array([v.bandof[0], v.bandof[1]])[typ]
    count = np.array([v.nband[0], v.nband[1]])[typ]
    assert np.all((band >= first) & (band < first + count))
    print("[gpu] cfg2: deposit", flush=True)
    node_q, _, cell_q, cell_n = eng.deposit()
    nc = m.ncell
    cell = (j.astype(np.int64) * m.cnx + i) * m.cnz + k
    np.testing.assert_array_equal(cell_n, np.bincount(cell, minlength=nc))
    q = p.charge[:p.n][live]
    fx = np.rint(q * FX_Q).astype(np.int64)
    assert int(cell_q[:nc].sum() + cell_q[nc:2 * nc].sum()) == int(fx.sum())
    assert int(cell_q[2 * nc:].sum()) == int(np.abs(fx).sum())
    # 8 corners rounded apiece: at most 4 units of 2^-48 per particle
    assert abs(int(node_q.sum()) - int(fx.sum())) <= 4 * n
    print("[gpu] cfg2: Poisson solve", flush=True)
    rhs = eng.rhs_from_charge(node_q)
    sol, it, rel = eng.poisson_solve(rhs, rtol=1e-10)
    assert rel <= 1e-10 and 0 < it < 100
    lib = oracle_lib.load()
    rhs_ref = np.zeros_like(rhs)
    lib.orc_rhs(C.byref(m), oracle_lib.i64ptr(node_q), None, oracle_lib.dptr(rhs_ref), None)
    ax = np.zeros_like(rhs)
    lib.orc_apply_full(C.byref(m), oracle_lib.dptr(sol), oracle_lib.dptr(ax))
    assert np.linalg.norm(ax - rhs_ref) <= 1e-8 * np.linalg.norm(rhs_ref)


def test_cfg2_steps_conserve_particles_and_charge(eng):
    eng.particles_init()
    p = eng.download()
    live = p.i[:p.n] >= 0
    n0, q0 = int(live.sum()), float(p.charge[:p.n][live].sum())
    for s in range(20):
        t = eng.step(1)
        d = emcgpu.tallies_dict(t)
        assert int(np.sum(d["drops"])) == 0, d["drops"]
        c = np.asarray(d["contact"], dtype=np.int64)
        p = eng.download()
        live = p.i[:p.n] >= 0
        n1, q1 = int(live.sum()), float(p.charge[:p.n][live].sum())
        assert n1 == n0 + int(t.n_gen) - int(t.n_catch), (s, n0, n1, t.n_gen, t.n_catch)
        dq = (c[:, 1].sum() - c[:, 4].sum()) / emcgpu.abi.FX_CONTACT
        tol = (int(t.n_gen) + int(t.n_catch) + 1) * 2.0 ** -40 + 1e-12 * np.abs(p.charge[:p.n][live]).sum()
        assert abs((q1 - q0) - dq) <= tol, (s, q1 - q0, dq)
        n0, q0 = n1, q1
    assert np.all(np.isfinite(eng.potential()))
    print(f"[gpu] cfg2: 20 steps, {n0} particles", flush=True)
