"""GPU parity: the HIP path against the oracle on identical inputs.

Bar: integer outputs (cell, tetrahedron, symmetry indices, counts, int64
tallies) are bit-exact; FP64 state (positions, wave vectors, energies) is
bit-exact as well (tolerance 0 ulp), which is stricter than the 1e-12
relative bound of the specification.
"""
import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu, field_like, sort_by_key, uniform_fields

pytestmark = pytest.mark.gpu

F64 = ("x", "y", "z", "kx", "ky", "kz", "energy", "charge")
I32 = ("i", "j", "k", "itet", "isym", "ptype")


def assert_same_population(pg, po, what=""):
    a, b = sort_by_key(pg), sort_by_key(po)
    assert a.n == b.n, f"{what}: live count GPU {a.n} vs oracle {b.n}"
    np.testing.assert_array_equal(a.key, b.key, err_msg=f"{what}: keys")
    for f in I32:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f"{what}: {f}")
    for f in F64:
        x, y = getattr(a, f), getattr(b, f)
        bad = np.nonzero(x != y)[0]
        assert bad.size == 0, f"{what}: {f} differs at {bad.size} particles, first {bad[:3]} {x[bad[:3]]} {y[bad[:3]]}"


def assert_same_tallies(tg, to, what=""):
    dg, do = emcgpu.tallies_dict(tg), emcgpu.tallies_dict(to)
    for k in dg:
        np.testing.assert_array_equal(np.asarray(dg[k]), np.asarray(do[k]), err_msg=f"{what}: tally {k}")


@pytest.fixture(scope="module")
def eng():
    print("[gpu] creating engine on the 20x40x12 FinFET deck", flush=True)
    e = emcgpu.Engine(cfg1_deck(), device=0)
    yield e
    e.close()


def test_flight_parity_multi_step(eng):
    """Five flight steps on a sampled population under a device-like field, heat on."""
    n = eng.particles_init()
    assert n > 20000
    pot, f = field_like(eng, seed=3, scale=20.0)
    eng.fields_set(f)
    eng.tallies_reset()
    p_or = eng.download()
    heat_o = np.zeros(eng.mesh.nnode, np.int64)
    rc_o = np.zeros(eng.mesh.pny, np.int64)
    for step in range(5):
        print(f"[gpu] flight parity step {step}", flush=True)
        tg = eng.fly_step(step, flag_heat=1)
        pg = eng.download()
        p_or, to, ex = oracle_lib.fly_step(eng, p_or, p_or.n, f, step, flag_heat=1, want_heat=True)
        heat_o += ex["heat_e"]
        rc_o += ex["rcur"]
        assert_same_tallies(tg, to, f"step {step}")
        assert_same_population(pg, p_or, f"step {step}")
        p_or = sort_by_key(p_or)
    he, hh = eng.heat()
    np.testing.assert_array_equal(he, heat_o)
    rc, lc = eng.currents()
    np.testing.assert_array_equal(rc, rc_o)


def test_flight_parity_hot_field(eng):
    """Strong uniform field: tetrahedron and BZ-face crossings, impact ionisation range."""
    eng.particles_init()
    f = uniform_fields(eng, ex=2.0, ey=30.0, ez=-3.0)
    eng.fields_set(f)
    p_or = eng.download()
    for step in range(3):
        print(f"[gpu] hot-field parity step {step}", flush=True)
        tg = eng.fly_step(100 + step)
        pg = eng.download()
        p_or, to, _ = oracle_lib.fly_step(eng, p_or, p_or.n, f, 100 + step)
        assert_same_tallies(tg, to, f"hot step {step}")
        assert_same_population(pg, p_or, f"hot step {step}")
        p_or = sort_by_key(p_or)
    assert tg.n_events > 2 * p_or.n // 2


def test_deposit_parity(eng):
    eng.particles_init()
    pot, f = field_like(eng, seed=5)
    eng.fields_set(f)
    eng.fly_step(7)
    p = eng.download()
    print("[gpu] deposit parity", flush=True)
    g = eng.deposit()
    o = oracle_lib.deposit(eng, p, p.n)
    for name, a, b in zip(("node_q", "node_stat", "cell_q", "cell_n"), g, o):
        np.testing.assert_array_equal(a, b, err_msg=name)
    live = p.i[:p.n] >= 0
    total = np.rint(p.charge[:p.n][live] * emcgpu.FX_Q).sum()
    assert abs(g[0].sum() - total) <= 8 * live.sum()
    # the deposit sorts the population by tile: same particles, empty slots gone
    assert not live.all(), "the flight step should have caught some particles"
    p2 = eng.download()
    assert p2.n == live.sum() and np.all(p2.i[:p2.n] >= 0)
    assert_same_population(p2, p.subset(np.nonzero(live)[0]), "tile sort")
    TI, TJ, TK = 4, 8, 4   # deposit tile, cells along i, j, k
    tile = ((p2.j[:p2.n] // TJ) * ((eng.mesh.cnx + TI - 1) // TI) + p2.i[:p2.n] // TI) * \
        ((eng.mesh.cnz + TK - 1) // TK) + p2.k[:p2.n] // TK
    assert np.all(np.diff(tile) >= 0), "population not in tile order"


def test_rhs_and_field_parity(eng):
    import ctypes as C
    lib = oracle_lib.load()
    eng.particles_init()
    node_q, _, cell_q, _ = eng.deposit()
    rhs_g = eng.rhs_from_charge(node_q)
    rhs_o = np.zeros(eng.mesh.nnode)
    lib.orc_rhs(C.byref(eng.mesh), oracle_lib.i64ptr(node_q), None, None, oracle_lib.dptr(rhs_o))
    np.testing.assert_array_equal(rhs_g, rhs_o)
    rho = np.zeros(eng.mesh.ncell)
    lib.orc_cell_rho(C.byref(eng.mesh), oracle_lib.i64ptr(cell_q), oracle_lib.dptr(rho))
    pot, _ = field_like(eng, seed=9)
    fg = eng.compute_field(pot, rho)
    fo = emcgpu.CellFieldArrays(eng.mesh.ncell)
    lib.orc_field(C.byref(eng.mesh), oracle_lib.dptr(pot), oracle_lib.dptr(fo.ex), oracle_lib.dptr(fo.ey),
                  oracle_lib.dptr(fo.ez))
    for k in ("ex", "ey", "ez"):
        np.testing.assert_array_equal(getattr(fg, k), getattr(fo, k), err_msg=k)
    lib.orc_surface_rates(C.byref(eng.mesh), oracle_lib.dptr(fo.ex), oracle_lib.dptr(fo.ey), oracle_lib.dptr(fo.ez),
                          oracle_lib.dptr(rho), oracle_lib.dptr(fo.e_rough), oracle_lib.dptr(fo.e_sph),
                          oracle_lib.dptr(fo.h_rough), oracle_lib.dptr(fo.h_sph))
    # pow() differs between the device library and glibc by at most a few ulp
    for k in ("e_rough", "e_sph", "h_rough", "h_sph"):
        np.testing.assert_allclose(getattr(fg, k), getattr(fo, k), rtol=1e-13, atol=0, err_msg=k)


def test_poisson_against_oracle(eng):
    import ctypes as C
    lib = oracle_lib.load()
    eng.particles_init()
    node_q, _, _, _ = eng.deposit()
    rhs = eng.rhs_from_charge(node_q)
    print("[gpu] poisson", flush=True)
    xg, it, rel = eng.poisson_solve(rhs, rtol=1e-11)
    assert rel <= 1e-11 and it > 0
    xo = np.zeros_like(rhs)
    rr = C.c_double(0)
    lib.orc_pcg(C.byref(eng.mesh), oracle_lib.dptr(rhs), oracle_lib.dptr(xo), 1e-12, 100000, C.byref(rr))
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
    # residual of the operator as assembled by init_matrix, with the contact rows
    rhs_ref = np.zeros_like(rhs)
    lib.orc_rhs(C.byref(eng.mesh), oracle_lib.i64ptr(node_q), None, oracle_lib.dptr(rhs_ref), None)
    ax = np.zeros_like(rhs)
    lib.orc_apply_full(C.byref(eng.mesh), oracle_lib.dptr(xg), oracle_lib.dptr(ax))
    assert np.linalg.norm(ax - rhs_ref) <= 1e-9 * np.linalg.norm(rhs_ref)


def test_full_step_composition(eng):
    """emc_step = deposit -> rhs -> PCG -> field -> fly; the flight of the step is
    bit-exact against the oracle fed with the step's own fields."""
    eng.particles_init()
    p0 = eng.download()
    eng.step(1)
    f = eng.fields_get()
    pot = eng.potential()
    pg = eng.download()
    # the step used step number 0 of this context's counter
    import ctypes as C
    lib = oracle_lib.load()
    node_q, _, cell_q, _ = oracle_lib.deposit(eng, p0, p0.n)
    rho = np.zeros(eng.mesh.ncell)
    lib.orc_cell_rho(C.byref(eng.mesh), oracle_lib.i64ptr(cell_q), oracle_lib.dptr(rho))
    np.testing.assert_array_equal(f.rho, rho)
    fo = emcgpu.CellFieldArrays(eng.mesh.ncell)
    lib.orc_field(C.byref(eng.mesh), oracle_lib.dptr(pot), oracle_lib.dptr(fo.ex), oracle_lib.dptr(fo.ey),
                  oracle_lib.dptr(fo.ez))
    np.testing.assert_array_equal(f.ex, fo.ex)
    po, to, _ = oracle_lib.fly_step(eng, p0, p0.n, f, 0)
    assert_same_population(pg, po, "full step")


def test_consecutive_steps_deposit(eng):
    """Twenty emc_step calls in a row: GENERATE clones fill holes and are appended,
    and each step's tile sort and deposit start from the previous step's order.
    Every step's cell densities equal the oracle's deposit of the population the
    step started from (bit-exact), and its potential equals a tight solve of the
    oracle's node charges."""
    import ctypes as C
    lib = oracle_lib.load()
    eng.particles_init()
    m = eng.mesh
    gen = 0
    for s in range(20):
        p = eng.download()
        t = eng.step(1)
        gen += t.n_gen
        print(f"[gpu] consecutive step {s}: {p.n} slots", flush=True)
        fg = eng.fields_get()
        node_q, _, cell_q, _ = oracle_lib.deposit(eng, p, p.n)
        rho = np.zeros(m.ncell)
        lib.orc_cell_rho(C.byref(m), oracle_lib.i64ptr(cell_q), oracle_lib.dptr(rho))
        np.testing.assert_array_equal(fg.rho, rho, err_msg=f"step {s}: cell densities")
        pg = eng.potential()
        x, it, rel = eng.poisson_solve(eng.rhs_from_charge(node_q), x0=pg, rtol=1e-12)
        assert np.max(np.abs(pg - x)) <= 1e-7 * np.max(np.abs(x)), f"step {s}: potential"
    assert gen > 0, "no GENERATE clones: the merge and append paths were not exercised"


def test_multiple_refresh_parity(eng):
    """emc_multiple_refresh against the oracle on one population, slot for slot and
    bit for bit: a flown population (holes included) with one particle of fifty
    times the charge in every third cell (those groups are resampled), some
    particles below pckill (dropped) and the rest kept."""
    eng.particles_init()
    pot, f = field_like(eng, seed=9)
    eng.fields_set(f)
    eng.fly_step(3)
    p = eng.download()
    m = eng.mesh
    live = np.nonzero(p.i[:p.n] >= 0)[0]
    c = (p.j[live].astype(np.int64) * m.cnx + p.i[live]) * m.cnz + p.k[live]
    g = 2 * c + p.ptype[live]
    first = np.unique(g, return_index=True)[1]
    p.charge[live[first[c[first] % 3 == 0]]] *= 50.0
    p.charge[live[::97]] = 1e-31
    eng.upload(p)
    print("[gpu] multiple refresh parity", flush=True)
    before, after = eng.multiple_refresh(29)
    pg = eng.download()
    po = oracle_lib.multiple_refresh(eng, p, p.n, 29)
    assert before == live.size
    assert after == po.n == pg.n
    for fld in emcgpu.PARTICLE_F64 + emcgpu.PARTICLE_I32 + ("key",):
        np.testing.assert_array_equal(getattr(pg, fld)[:pg.n], getattr(po, fld)[:po.n], err_msg=fld)
    assert not np.all(np.isin(pg.key[:pg.n], p.key[:p.n])), "no group was resampled"


def test_device_initial_sampling_equals_host(eng):
    """emc_particles_init on the GPU (k_sample_population) gives the host
    sampler's population slot for slot and bit for bit (same code, same draws)."""
    print("[gpu] device initial sampling", flush=True)
    eng.particles_init()
    pg = eng.download()
    h = emcgpu.Engine(cfg1_deck(), device=-1)
    try:
        h.particles_init()
        ph = h.download()
    finally:
        h.close()
    assert pg.n == ph.n
    for fld in emcgpu.PARTICLE_F64 + emcgpu.PARTICLE_I32 + ("key",):
        np.testing.assert_array_equal(getattr(pg, fld)[:pg.n], getattr(ph, fld)[:ph.n], err_msg=fld)


def test_checkpoint_resume(tmp_path):
    """Save after three steps, then four more steps from the saved context and from
    a fresh context that loaded the file: same particles, tallies and potential."""
    path = str(tmp_path / "emc.ckpt")
    a = emcgpu.Engine(cfg1_deck(), device=0)
    b = emcgpu.Engine(cfg1_deck(), device=0)
    try:
        a.particles_init()
        a.step(3)
        a.checkpoint_save(path)
        print("[gpu] checkpoint saved, stepping both contexts", flush=True)
        ta = a.step(4)
        b.checkpoint_load(path)
        tb = b.step(4)
        assert_same_population(a.download(), b.download(), "resume")
        assert_same_tallies(ta, tb, "resume")
        np.testing.assert_array_equal(a.potential(), b.potential())
        for x, y in zip(a.heat(), b.heat()):
            np.testing.assert_array_equal(x, y)
        for x, y in zip(a.currents(), b.currents()):
            np.testing.assert_array_equal(x, y)
    finally:
        a.close()
        b.close()


def test_poisson_multilevel_mesh():
    """60x120x36 cells (273 k nodes): the hierarchy has two grid-wide levels before
    the one-workgroup bottom, so every V-cycle kernel runs.  Random charge on the
    unknown nodes, contact potentials from the deck; checked against the oracle CG
    and the assembled operator."""
    import ctypes as C
    from helpers import deck
    lib = oracle_lib.load()
    print("[gpu] poisson on the 60x120x36 mesh", flush=True)
    e = emcgpu.Engine(deck("mg3", refine=3, electrons=1000, holes=100, vd=0.5, vg=0.7), device=0)
    try:
        levels = e.mg_levels()
        assert len(levels) >= 4 and np.prod(levels[1][0]) > 16384
        m = e.mesh
        rng = np.random.default_rng(11)
        node_q = (rng.standard_normal(m.nnode) * 1e3 * emcgpu.FX_Q).astype(np.int64)
        rhs = e.rhs_from_charge(node_q)
        xg, it, rel = e.poisson_solve(rhs, rtol=1e-11)
        assert rel <= 1e-11 and 0 < it <= 30, (it, rel)
        xo = np.zeros_like(rhs)
        rr = C.c_double(0)
        lib.orc_pcg(C.byref(m), oracle_lib.dptr(rhs), oracle_lib.dptr(xo), 1e-12, 100000, C.byref(rr))
        assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
        rhs_ref = np.zeros_like(rhs)
        lib.orc_rhs(C.byref(m), oracle_lib.i64ptr(node_q), None, oracle_lib.dptr(rhs_ref), None)
        ax = np.zeros_like(rhs)
        lib.orc_apply_full(C.byref(m), oracle_lib.dptr(xg), oracle_lib.dptr(ax))
        assert np.linalg.norm(ax - rhs_ref) <= 1e-9 * np.linalg.norm(rhs_ref)
        # warm start from the solution: at most one more iteration
        x2, it2, rel2 = e.poisson_solve(rhs, x0=xg, rtol=1e-11)
        assert it2 <= 1 and rel2 <= 1e-11
        assert np.linalg.norm(x2 - xg) <= 1e-10 * np.linalg.norm(xg)
    finally:
        e.close()


def test_flight_parity_large_population():
    """128 k particles in a strong random field, five steps with heat: many lanes of
    a wave scatter in the same step, which a flight kernel that mixes per-lane
    state across lanes gets wrong (the rejection loop of get_state is the case
    this caught)."""
    from helpers import deck
    print("[gpu] flight parity, 128k particles", flush=True)
    e = emcgpu.Engine(deck("big", refine=1, electrons=100000, holes=20000, vd=0.5, vg=0.7), device=0)
    try:
        e.particles_init()
        _, f = field_like(e, seed=7, scale=60.0)
        e.fields_set(f)
        e.tallies_reset()
        po = e.download()
        for step in range(5):
            tg = e.fly_step(50 + step, flag_heat=1)
            pg = e.download()
            po, to, _ = oracle_lib.fly_step(e, po, po.n, f, 50 + step, flag_heat=1)
            assert_same_tallies(tg, to, f"large step {step}")
            assert_same_population(pg, po, f"large step {step}")
            po = sort_by_key(po)
        assert tg.n_full_loop > 0 and tg.n_phonon + tg.n_impurity > 100
    finally:
        e.close()


def test_flight_parity_near_band_edge(eng):
    """Electrons placed on the iso-energy surfaces just above the intervalley phonon
    energies (20.5, 48.9 and 62.7 meV; their wave vectors and tetrahedra from
    emc_band_final_state), in a weak field, twenty steps: phonon emission leaves many
    of them 1-3 meV above the band edge, where get_state's rejection loop takes
    hundreds to thousands of tries (at 1.6 meV about one search in five reaches the
    10 000-try limit), while the other lanes of the wave take their own paths.  The
    slow kernel takes those searches with the whole wave, 128 draws at a time after
    the particle's first try.  Population and tallies equal the oracle's bit
    for bit."""
    eng.particles_init()
    p = eng.download()
    n = p.n
    el = np.nonzero((p.i[:n] >= 0) & (p.ptype[:n] == 0))[0]
    rng = np.random.default_rng(21)
    targets = np.array([0.0205, 0.0489, 0.0627])
    energy = targets[rng.integers(0, len(targets), len(el))]
    isym = rng.integers(0, 48, len(el)).astype(np.int32)
    band = np.full(len(el), eng.band.bandof[0], np.int32)
    k, itet, nd, sc = eng.band_final_state(p.key[el], 12345, energy, band, isym, np.zeros(len(el), np.int32))
    ok = sc == 0
    q = el[ok]
    p.kx[q], p.ky[q], p.kz[q] = k[ok, 0], k[ok, 1], k[ok, 2]
    p.energy[q] = energy[ok]
    p.itet[q] = itet[ok]
    p.isym[q] = isym[ok]
    assert ok.sum() > 0.9 * len(el)
    # the final states these emissions reach take long loops
    kk, _, nd2, _ = eng.band_final_state(p.key[el[:400]], 7, np.full(400, 0.0015), band[:400], isym[:400],
                                         np.zeros(400, np.int32))
    assert np.median(nd2) > 100
    eng.upload(p)
    f = uniform_fields(eng, ex=0.5, ey=2.0, ez=-0.5)
    eng.fields_set(f)
    eng.tallies_reset()
    po = eng.download()
    phonons = 0
    for step in range(20):
        print(f"[gpu] near-band-edge parity step {step}", flush=True)
        tg = eng.fly_step(300 + step)
        pg = eng.download()
        po, to, _ = oracle_lib.fly_step(eng, po, po.n, f, 300 + step)
        assert_same_tallies(tg, to, f"near-edge step {step}")
        assert_same_population(pg, po, f"near-edge step {step}")
        po = sort_by_key(po)
        phonons += tg.n_phonon
    assert phonons > 100


def test_multiple_refresh_grows_the_population():
    """A refresh whose output needs more slots than the population was given:
    the buffers grow (new particle buffers, flight lists and work arrays), the
    output equals the oracle's slot for slot, and the engine steps and refreshes
    on from the grown buffers."""
    h = emcgpu.Engine(cfg1_deck(), device=-1)
    try:
        h.particles_init()
        full = h.download()
    finally:
        h.close()
    m_e = emcgpu.Engine(cfg1_deck(), device=0)
    try:
        m = m_e.mesh
        live = np.nonzero(full.i[:full.n] >= 0)[0]
        c = (full.j[live].astype(np.int64) * m.cnx + full.i[live]) * m.cnz + full.k[live]
        g = 2 * c + full.ptype[live]
        first = live[np.unique(g, return_index=True)[1]][:3000]
        p = full.subset(first)
        p.n = len(first)
        m_e.upload(p)
        print(f"[gpu] refresh growth: {p.n} particles uploaded", flush=True)
        before, after = m_e.multiple_refresh(5)
        po = oracle_lib.multiple_refresh(m_e, p, p.n, 5)
        assert before == p.n and after == po.n
        assert after > 1.25 * p.n + 1024, "the refresh did not need more slots"
        pg = m_e.download()
        for fld in emcgpu.PARTICLE_F64 + emcgpu.PARTICLE_I32 + ("key",):
            np.testing.assert_array_equal(getattr(pg, fld)[:pg.n], getattr(po, fld)[:po.n], err_msg=fld)
        m_e.step(2)
        q = m_e.download()
        before2, after2 = m_e.multiple_refresh(9)
        qo = oracle_lib.multiple_refresh(m_e, q, q.n, 9)
        qg = m_e.download()
        assert after2 == qo.n
        for fld in emcgpu.PARTICLE_F64 + emcgpu.PARTICLE_I32 + ("key",):
            np.testing.assert_array_equal(getattr(qg, fld)[:qg.n], getattr(qo, fld)[:qo.n], err_msg=fld)
    finally:
        m_e.close()
