"""Pin the oracle flight (particle_fly and callees) by invariants that a
plausible mistake breaks: conservation, drift/energy bookkeeping, the particle
staying inside its tetrahedron and cell, the BZ confinement, closed-form exit
times and order independence."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu, field_like, sort_by_key, uniform_fields


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(), device=-1)
    e.particles_init()
    e.pop0 = e.download()
    e.tabs = e.band_arrays()
    e.marr = e.mesh_arrays()
    return e


def in_wedge(b, isym, k):
    ms = b["matsym"].reshape(6, 48)
    out = np.zeros_like(k)
    for r in range(3):
        out[np.arange(len(isym)), ms[r, isym] - 1] = k[:, r] * ms[3 + r, isym]
    return out


def lin_energy(b, itet, isym, k):
    nt = len(b["ibt"])
    kw = in_wedge(b, isym, k)
    c0 = b["tet"].reshape(4, nt)[0, itet]
    vt = b["vt"].reshape(nt, 4)[itet]
    d = kw - np.stack([b["xkk"][c0], b["ykk"][c0], b["zkk"][c0]], axis=1)
    return vt[:, 3] + vt[:, 0] * d[:, 0] + vt[:, 1] * d[:, 1] + vt[:, 2] * d[:, 2]


def containment(b, itet, isym, k):
    nt = len(b["ibt"])
    kw = in_wedge(b, isym, k)
    dnt = b["dnt"].reshape(4, 4, nt)
    return np.stack([dnt[0, f, itet] * kw[:, 0] + dnt[1, f, itet] * kw[:, 1] + dnt[2, f, itet] * kw[:, 2]
                     - dnt[3, f, itet] for f in range(4)], axis=1)


def check_state(e, p, tol_e=1e-9):
    b, m = e.tabs, e.marr
    live = p.i[:p.n] >= 0
    q = p.subset(np.nonzero(live)[0])
    k = np.stack([q.kx, q.ky, q.kz], axis=1)
    # inside the containing tetrahedron (tolerance scaled like MY_ZERO with the normal length)
    d = containment(b, q.itet, q.isym, k)
    assert np.all(d >= -1e-9), f"{np.sum(d < -1e-9)} particles outside their tetrahedron"
    # energy equals the tetrahedron's linear dispersion at k
    el = lin_energy(b, q.itet, q.isym, k)
    assert np.max(np.abs(el - q.energy)) < tol_e * max(1.0, np.max(np.abs(q.energy)))
    # band of the tetrahedron matches the particle type
    ibt = b["ibt"][q.itet]
    assert np.all((ibt < 4) == (q.ptype == 0))
    # confined to the first Brillouin zone (CONFBZ, src/Band.cpp:573-616)
    a0 = e.band.a0pi
    assert np.all(np.abs(k) <= a0 * (1 + 1e-12))
    assert np.all(np.abs(k).sum(axis=1) <= 1.5 * a0 * (1 + 1e-12))
    # inside its cell (BETWEEN01 with MY_ZERO, include/defs.h:128)
    for ax, idx in (("x", q.i), ("y", q.j), ("z", q.k)):
        l, dl = m["l" + ax], m["d" + ax]
        r = (getattr(q, ax) - l[idx]) / dl[idx]
        assert np.all((r >= -1e-9) & (r <= 1 + 1e-9)), ax
    # carriers only in silicon
    cell = (q.j * e.mesh.cnx + q.i) * e.mesh.cnz + q.k
    assert np.all(m["material"][cell] == 3)


def test_initial_population_is_valid(eng):
    check_state(eng, eng.pop0)
    ev = eng.pop0.energy * eng.mesh.eV0
    assert 0.030 < ev.mean() < 0.050  # ~3/2 kT at 300 K


def test_conservation_and_state_after_steps(eng):
    pot, f = field_like(eng, seed=11, scale=15.0)
    p = eng.pop0
    for s in range(6):
        live_in = int(np.sum(p.i[:p.n] >= 0))
        p, t, _ = oracle_lib.fly_step(eng, p, p.n, f, s)
        live_out = int(np.sum(p.i[:p.n] >= 0))
        # particle conservation (check_par_number, src/mcmodel.cpp:1414-1455)
        assert live_out == live_in - t.n_catch + t.n_gen - t.drops[8]
        assert sum(t.drops[:6]) == 0
        check_state(eng, p)
        p = sort_by_key(p)


def test_drift_without_scattering_is_ballistic(eng):
    """One step in a uniform field: particles that stayed in the same tetrahedron and
    cell and did not scatter moved by k += F dt, E += v.F dt, r += v dt exactly."""
    f = uniform_fields(eng, ex=1.0, ey=-2.0, ez=0.5)
    p0 = eng.pop0
    p1, t, _ = oracle_lib.fly_step(eng, p0, p0.n, f, 3)
    b = eng.tabs
    dt = eng.mesh.dt
    same = (p1.i[:p0.n] >= 0) & (p1.itet[:p0.n] == p0.itet) & (p1.isym[:p0.n] == p0.isym) & \
           (p1.i[:p0.n] == p0.i) & (p1.j[:p0.n] == p0.j) & (p1.k[:p0.n] == p0.k)
    idx = np.nonzero(same)[0]
    sign = np.where(p0.ptype[idx] == 0, -1.0, 1.0)
    F = np.stack([sign * 1.0, sign * -2.0, sign * 0.5], axis=1)
    k0 = np.stack([p0.kx[idx], p0.ky[idx], p0.kz[idx]], axis=1)
    k1 = np.stack([p1.kx[idx], p1.ky[idx], p1.kz[idx]], axis=1)
    moved = np.abs(k1 - k0 - F * dt).max(axis=1) < 1e-12
    # particles with a real scattering event changed k by more than the drift
    assert moved.mean() > 0.97
    sel = idx[moved]
    nt = len(b["ibt"])
    vt = b["vt"].reshape(nt, 4)[p0.itet[sel]]
    ms = b["matsym"].reshape(6, 48)
    v = np.zeros((len(sel), 3))
    for r in range(3):
        v[:, r] = vt[np.arange(len(sel)), ms[r, p0.isym[sel]] - 1] * ms[3 + r, p0.isym[sel]]
    Fs = F[moved]
    dE = (v * Fs).sum(axis=1) * dt
    np.testing.assert_allclose(p1.energy[sel] - p0.energy[sel], dE, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(p1.y[sel] - p0.y[sel], v[:, 1] * dt, rtol=1e-9, atol=1e-18)


def test_tet_time_closed_form(eng):
    """TetTime (src/TetHit.cpp:4-106): k + F t exits through face itetdir at t, and is
    inside the tetrahedron before."""
    lib = oracle_lib.load()
    b = eng.tabs
    nt = len(b["ibt"])
    rng = np.random.default_rng(5)
    P = [np.stack([b["xkk"][b["tet"].reshape(4, nt)[c]], b["ykk"][b["tet"].reshape(4, nt)[c]],
                   b["zkk"][b["tet"].reshape(4, nt)[c]]], axis=1) for c in range(4)]
    ms = b["matsym"].reshape(6, 48)
    n_checked = 0
    for _ in range(400):
        it = int(rng.integers(0, nt))
        isym = int(rng.integers(0, 48))
        w = rng.dirichlet(np.ones(4))
        kw = sum(w[c] * P[c][it] for c in range(4))
        k = np.zeros(3)
        for r in range(3):
            k[r] = kw[ms[r, isym] - 1] * ms[3 + r, isym]
        F = rng.normal(size=3)
        d = C.c_int(0)
        t = lib.orc_tet_time(C.byref(eng.band), it, isym, k.ctypes.data_as(C.POINTER(C.c_double)),
                             F.ctypes.data_as(C.POINTER(C.c_double)), C.byref(d))
        kt = (k + F * t)[None, :]
        dd = containment(b, np.array([it]), np.array([isym]), kt)[0]
        scale = np.abs(b["dnt"].reshape(4, 4, nt)[:3, d.value, it]).sum() * eng.band.a0pi
        assert abs(dd[d.value]) < 1e-12 * scale
        assert np.all(dd > -1e-12 * scale)
        half = containment(b, np.array([it]), np.array([isym]), (k + F * t * 0.5)[None, :])[0]
        assert np.all(half > -1e-12 * scale)
        n_checked += 1
    assert n_checked == 400


def test_get_state_lands_on_surface(eng):
    """get_state (src/ParticleUtils.cpp:48-120): the final k lies in the chosen
    tetrahedron of the requested band, on the iso-energy surface."""
    lib = oracle_lib.load()
    b = eng.tabs
    rng = np.random.default_rng(6)
    hits = 0
    for trial in range(300):
        band = int(rng.integers(0, 7))
        E = float(rng.uniform(0.2, 60.0))
        isym = int(rng.integers(0, 48))
        k = np.zeros(3)
        it = C.c_int(0)
        nd = C.c_int(0)
        self_sc = lib.orc_get_state(C.byref(eng.band), trial, 1, E, band, isym,
                                    k.ctypes.data_as(C.POINTER(C.c_double)), C.byref(it), C.byref(nd))
        if self_sc:
            continue
        hits += 1
        assert b["ibt"][it.value] == band
        d = containment(b, np.array([it.value]), np.array([isym]), k[None, :])[0]
        assert np.all(d > -1e-9)
        el = lin_energy(b, np.array([it.value]), np.array([isym]), k[None, :])[0]
        assert abs(el - E) < 1e-9 * max(1, E)
        assert nd.value >= 4
    assert hits > 150


def test_getisym_finds_containing_tetrahedron(eng):
    """Getisym (src/ParticleUtils.cpp:395-625) on points of random tetrahedra."""
    lib = oracle_lib.load()
    b = eng.tabs
    nt = len(b["ibt"])
    rng = np.random.default_rng(7)
    tet = b["tet"].reshape(4, nt)
    ms = b["matsym"].reshape(6, 48)
    for _ in range(200):
        it = int(rng.integers(0, nt))
        w = rng.dirichlet(np.ones(4) * 2)
        kw = sum(w[c] * np.array([b["xkk"][tet[c, it]], b["ykk"][tet[c, it]], b["zkk"][tet[c, it]]]) for c in range(4))
        s = int(rng.integers(0, 48))
        k = np.array([kw[ms[r, s] - 1] * ms[3 + r, s] for r in range(3)])
        E = float(lin_energy(b, np.array([it]), np.array([s]), k[None, :])[0])
        isym, itet = C.c_int(0), C.c_int(0)
        st = lib.orc_getisym(C.byref(eng.band), k.ctypes.data_as(C.POINTER(C.c_double)), E, int(b["ibt"][it]), 0,
                             C.byref(isym), C.byref(itet))
        assert st == 0
        d = containment(b, np.array([itet.value]), np.array([isym.value]), k[None, :])[0]
        assert np.all(d >= -1e-13 * 1e3)
        assert b["ibt"][itet.value] == b["ibt"][it]


def test_order_independence_and_determinism(eng):
    """Per-particle results and all int64 tallies do not depend on particle order."""
    pot, f = field_like(eng, seed=12, scale=15.0)
    p0 = eng.pop0
    a, ta, xa = oracle_lib.fly_step(eng, p0, p0.n, f, 9, flag_heat=1, want_heat=True)
    perm = np.random.default_rng(8).permutation(p0.n)
    b, tb, xb = oracle_lib.fly_step(eng, p0.subset(perm), p0.n, f, 9, flag_heat=1, want_heat=True)
    sa, sb = sort_by_key(a), sort_by_key(b)
    for fld in ("x", "y", "z", "kx", "ky", "kz", "energy", "i", "j", "k", "itet", "isym"):
        np.testing.assert_array_equal(getattr(sa, fld), getattr(sb, fld))
    da, db = emcgpu.tallies_dict(ta), emcgpu.tallies_dict(tb)
    for kk in da:
        np.testing.assert_array_equal(np.asarray(da[kk]), np.asarray(db[kk]))
    for kk in xa:
        np.testing.assert_array_equal(xa[kk], xb[kk])


def test_contacts_conserve_particles_and_charge(eng):
    """Every step conserves particles and charge through the contact tallies:
    live count after = before + generated - caught, the per-contact counts add
    up to the step totals, and the summed charge moves by exactly the
    generated minus the caught charge (each tally term is rounded to 2^-40).
    A GENERATE that forgot its tally, a CATCH that left the particle alive or
    a clone with the wrong charge fails it.  (Ten 0.1 fs steps move a carrier
    about 0.1 nm, so the contact events are thermal exchange and the sign of
    the net current is noise at this length; the whole-run current is checked
    on the GPU against the reference's own windows.)"""
    pot, f = field_like(eng, seed=13, scale=20.0)
    p = eng.pop0
    ngen = ncat = 0
    for s in range(10):
        live0 = p.i[:p.n] >= 0
        n0, q0 = int(live0.sum()), float(p.charge[:p.n][live0].sum())
        p, t, _ = oracle_lib.fly_step(eng, p, p.n, f, 50 + s)
        d = emcgpu.tallies_dict(t)
        c = np.asarray(d["contact"], dtype=np.int64)
        assert int(np.sum(d["drops"][:6])) == 0
        live1 = p.i[:p.n] >= 0
        n1, q1 = int(live1.sum()), float(p.charge[:p.n][live1].sum())
        assert n1 == n0 + int(t.n_gen) - int(t.n_catch)
        assert int(c[:, 0].sum()) == int(t.n_gen)
        assert int(c[:, 3].sum()) == int(t.n_catch)
        dq = (c[:, 1].sum() - c[:, 4].sum()) / emcgpu.abi.FX_CONTACT
        tol = (int(t.n_gen) + int(t.n_catch) + 1) * 2.0 ** -40 + 1e-12 * np.abs(p.charge[:p.n][live1]).sum()
        assert abs((q1 - q0) - dq) <= tol, (q1 - q0, dq)
        ngen += int(t.n_gen)
        ncat += int(t.n_catch)
        p = sort_by_key(p)
    assert ngen > 0 and ncat > 0
