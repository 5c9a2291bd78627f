"""GPU: the quantum-correction effective potential (src/mcmodel.cpp:2337-2405) and
the step with it on (compute_field, src/mcmodel.cpp:2239-2263).

emc_qc_potential equals the oracle's restatement bit for bit.  From step
quantum_start_step on, emc_step gives electrons the field of the corrected
potential and holes the field of the Poisson potential, and the flight under
those fields is the oracle's; the window keeps both potential sums.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu
from test_gpu_parity import assert_same_population

pytestmark = pytest.mark.gpu


def oracle_qc(eng, pot):
    out = np.zeros(eng.mesh.nnode)
    lib = oracle_lib.load()
    assert lib.orc_qc_potential(C.byref(eng.mesh), oracle_lib.dptr(np.ascontiguousarray(pot)),
                                oracle_lib.dptr(out)) == 0
    return out


def oracle_field(eng, pot):
    lib = oracle_lib.load()
    f = [np.zeros(eng.mesh.ncell) for _ in range(3)]
    lib.orc_field(C.byref(eng.mesh), oracle_lib.dptr(np.ascontiguousarray(pot)), *[oracle_lib.dptr(x) for x in f])
    return f


def test_qc_potential_and_step():
    e = emcgpu.Engine(cfg1_deck(qc=1, quantum_start_step=2), device=0)
    try:
        m = e.mesh
        rng = np.random.default_rng(11)
        for what, pot in (("random", rng.normal(size=m.nnode)), ("smooth", np.sin(np.arange(m.nnode) * 1e-3))):
            print(f"[gpu] qc potential, {what} input", flush=True)
            np.testing.assert_array_equal(e.qc_potential(pot), oracle_qc(e, pot), err_msg=what)
        e.particles_init()
        e.tallies_reset()
        print("[gpu] qc steps 0-1 (correction not yet active)", flush=True)
        e.step(1)
        p1 = e.potential()
        e.step(1)
        p2 = e.potential()
        f = e.fields_get()
        np.testing.assert_array_equal(f.ex, f.hx)
        p0 = e.download()
        print("[gpu] qc step 2 (active)", flush=True)
        e.step(1)
        pot = e.potential()
        f = e.fields_get()
        qpot = oracle_qc(e, pot)
        for got, want, name in zip((f.ex, f.ey, f.ez), oracle_field(e, qpot), ("ex", "ey", "ez")):
            np.testing.assert_array_equal(got, want, err_msg="electron " + name)
        for got, want, name in zip((f.hx, f.hy, f.hz), oracle_field(e, pot), ("hx", "hy", "hz")):
            np.testing.assert_array_equal(got, want, err_msg="hole " + name)
        assert np.any(f.ex != f.hx)
        po, _, _ = oracle_lib.fly_step(e, p0, p0.n, f, 2)
        assert_same_population(e.download(), po, "qc step")
        # the window: statistic() adds the potentials each step starts from (the
        # initial zeros, then each solve); the corrected sum takes the corrected
        # potential from the step after the first corrected solve on
        print("[gpu] qc step 3 (window sums)", flush=True)
        e.step(1)
        w = e.stat_window()
        assert w["steps"] == 4
        z = np.zeros(m.nnode)
        np.testing.assert_array_equal(w["stat_pot"], ((z + p1) + p2) + pot)
        np.testing.assert_array_equal(w["stat_qc_pot"], ((z + p1) + p2) + qpot)
    finally:
        e.close()


def test_dispersive_phonon_flight():
    """With the quantum correction on, EPSCAT takes the g-LO and f-LA/TO final
    energies from the dispersive tables (src/ElectronPhononScatter.cpp:148-163):
    two flight steps of 1e-14 s under a device-like field (thousands of phonon
    events) equal the oracle's, particles and tallies bit for bit."""
    from helpers import field_like
    from test_gpu_parity import assert_same_tallies
    e = emcgpu.Engine(cfg1_deck(qc=1, dt=1e-14), device=0)
    try:
        e.particles_init()
        pot, f = field_like(e, seed=5, scale=20.0)
        e.fields_set(f)
        p = e.download()
        phonon = 0
        for step in range(2):
            print(f"[gpu] dispersive-table flight step {step}", flush=True)
            tg = e.fly_step(step)
            p, to, _ = oracle_lib.fly_step(e, p, p.n, f, step)
            assert_same_population(e.download(), p, f"step {step}")
            assert_same_tallies(tg, to, f"step {step}")
            phonon += tg.n_phonon
        assert phonon > 1000, phonon
    finally:
        e.close()


def test_checkpoint_resume_with_correction_and_heat_phase(tmp_path):
    """A checkpoint taken with the quantum correction active, and one taken in the
    heating phase, resume exactly: particles, potentials, both potential sums of
    the window, heat."""
    deck = cfg1_deck(qc=1, quantum_start_step=1, stat_step=4)
    a = emcgpu.Engine(deck, device=0)
    b = emcgpu.Engine(deck, device=0)
    try:
        a.particles_init()
        a.tallies_reset()
        a.step(3)
        path = str(tmp_path / "qc.ckpt")
        a.checkpoint_save(path)
        b.checkpoint_load(path)
        print("[gpu] resume with the correction active", flush=True)
        a.step(2)
        b.step(2)
        assert_same_population(a.download(), b.download(), "qc resume")
        wa, wb = a.stat_window(), b.stat_window()
        for k in ("stat_pot", "stat_qc_pot", "node_stat", "cell_count"):
            np.testing.assert_array_equal(wa[k], wb[k], err_msg=k)
        np.testing.assert_array_equal(a.fields_get().ex, b.fields_get().ex)
        # heating phase: window means of the last window, then heat steps
        a.write_stat(wa, str(tmp_path), 4)
        a.tallies_reset()
        a.heat_begin()
        a.heat_step(1)
        path = str(tmp_path / "heat.ckpt")
        a.checkpoint_save(path)
        b.checkpoint_load(path)
        print("[gpu] resume in the heating phase", flush=True)
        a.heat_step(2)
        b.heat_step(2)
        assert_same_population(a.download(), b.download(), "heat resume")
        for x, y in zip(a.heat(), b.heat()):
            np.testing.assert_array_equal(x, y)
        wa, wb = a.stat_window(), b.stat_window()
        for k in ("stat_pot", "stat_qc_pot", "node_stat", "heat_e", "heat_h"):
            np.testing.assert_array_equal(wa[k], wb[k], err_msg=k)
        fa, fb = a.fields_get(), b.fields_get()
        np.testing.assert_array_equal(fa.ex, fb.ex)
        np.testing.assert_array_equal(fa.hx, fb.hx)
    finally:
        a.close()
        b.close()
