"""GPU: the statistics window of emc_step and the heating phase.

The window sums (statistic(), src/mcmodel.cpp:1167-1408) equal the sums of
their parts taken step by step: the oracle's deposit statistics of the
population each step starts from, the potential each step starts from (added
in the same order, so bit for bit), the per-cell counts, and the step tallies.
One heat step (heat_density, src/heat.cpp:48-98) equals its parts: the oracle's
cell densities of the step's population, the field of the fixed window-mean
potential, and the oracle's flight with heat tallies under those fields.
"""
import ctypes as C
import os
import filecmp

import numpy as np
import pytest

import oracle_lib
from helpers import REPO, cfg1_deck, emcgpu
from test_gpu_parity import assert_same_population, assert_same_tallies

pytestmark = pytest.mark.gpu

STAT_STEP = 4
HEAT_STEP0 = 1 << 40


def add_tallies(acc, t):
    for c in range(emcgpu.abi.MAX_CONTACT):
        for k in range(emcgpu.abi.CT_NCOL):
            acc.contact[c][k] += t.contact[c][k]
    for r in range(emcgpu.abi.NREASON):
        acc.drops[r] += t.drops[r]
    for f in ("n_catch", "n_gen", "n_phonon", "n_impurity", "n_events"):
        setattr(acc, f, getattr(acc, f) + getattr(t, f))


def test_window_sums_and_files(tmp_path):
    e = emcgpu.Engine(cfg1_deck(stat_step=STAT_STEP), device=0)
    try:
        m = e.mesh
        e.particles_init()
        e.tallies_reset()
        pot_sum = np.zeros(m.nnode)
        stat = np.zeros(10 * m.nnode, np.int64)
        count = np.zeros(m.ncell, np.int64)
        acc = emcgpu.abi.FlyTallies()
        for s in range(6):
            print(f"[gpu] window step {s}", flush=True)
            p = e.download()
            pot_sum += e.potential()
            _, ns, _, cn = oracle_lib.deposit(e, p, p.n)
            stat += ns
            count += cn
            add_tallies(acc, e.step(1))
        w = e.stat_window()
        assert w["steps"] == 6
        np.testing.assert_array_equal(w["stat_pot"], pot_sum)
        np.testing.assert_array_equal(w["node_stat"], stat)
        np.testing.assert_array_equal(w["cell_count"], count)
        assert_same_tallies(w["tallies"], acc, "window")
        assert not w["heat_e"].any() and not w["heat_h"].any(), "the transient tallies no heat"
        rc, lc = e.currents()
        np.testing.assert_array_equal(w["rcur"], rc)
        np.testing.assert_array_equal(w["lcur"], lc)
        # emc_output_stat writes what emc_write_stat writes from the same window
        (tmp_path / "a").mkdir()
        (tmp_path / "b").mkdir()
        e.output_stat(str(tmp_path / "a"), 5)
        e.write_stat(w, str(tmp_path / "b"), 5)
        names = ["current", "section_current5", "electron_num5", "pot5", "qc_pot5", "Electron5", "Hole5"]
        match, mismatch, errors = filecmp.cmpfiles(tmp_path / "a", tmp_path / "b", names, shallow=False)
        assert match == names, (mismatch, errors)
        e.output_for_restart(str(tmp_path / "a"), 5)
        rows = (tmp_path / "a" / "par_info_restart5").read_text().split("\n")
        assert rows[-2] == "-1 -1 -1" and len(rows) == m.ncell + 2
        p = e.download()
        assert sum(int(r.split()[4]) + int(r.split()[6]) for r in rows[:-2]) == int((p.i[:p.n] >= 0).sum())
        # a reset starts an empty window
        e.tallies_reset()
        w = e.stat_window()
        assert w["steps"] == 0 and not w["stat_pot"].any() and not w["node_stat"].any()
        assert not w["cell_count"].any() and w["tallies"].n_phonon == 0
    finally:
        e.close()


def test_heat_phase_step(tmp_path):
    lib = oracle_lib.load()
    e = emcgpu.Engine(cfg1_deck(stat_step=STAT_STEP), device=0)
    try:
        m = e.mesh
        e.particles_init()
        e.tallies_reset()
        print("[gpu] heat phase: transient window", flush=True)
        e.step(STAT_STEP)
        w = e.stat_window()
        e.output_stat(str(tmp_path), STAT_STEP - 1)
        pot_avg = w["stat_pot"] / STAT_STEP
        pot_last = e.potential()
        e.tallies_reset()
        e.heat_begin()
        p0 = e.download()
        print("[gpu] heat phase: first heat step", flush=True)
        e.heat_step(1)
        f = e.fields_get()
        pg = e.download()
        h0 = e.heat()
        _, ns0, cell_q, _ = oracle_lib.deposit(e, p0, p0.n)
        rho = np.zeros(m.ncell)
        lib.orc_cell_rho(C.byref(m), oracle_lib.i64ptr(cell_q), oracle_lib.dptr(rho))
        np.testing.assert_array_equal(f.rho, rho)
        fo = emcgpu.CellFieldArrays(m.ncell)
        lib.orc_field(C.byref(m), oracle_lib.dptr(pot_avg), oracle_lib.dptr(fo.ex), oracle_lib.dptr(fo.ey),
                      oracle_lib.dptr(fo.ez))
        for k in ("ex", "ey", "ez"):
            np.testing.assert_array_equal(getattr(f, k), getattr(fo, k), err_msg=k)
        po, to, ex = oracle_lib.fly_step(e, p0, p0.n, f, HEAT_STEP0, flag_heat=1, want_heat=True)
        assert_same_population(pg, po, "heat step")
        np.testing.assert_array_equal(h0[0], ex["heat_e"])
        np.testing.assert_array_equal(h0[1], ex["heat_h"])
        assert h0[0].any(), "no heat tallied"
        print("[gpu] heat phase: second heat step", flush=True)
        _, ns1, _, _ = oracle_lib.deposit(e, pg, pg.n)
        e.heat_step(1)
        w2 = e.stat_window()
        assert w2["steps"] == 2
        np.testing.assert_array_equal(w2["heat_e"], h0[0])
        np.testing.assert_array_equal(w2["heat_h"], h0[1])
        np.testing.assert_array_equal(w2["stat_pot"], pot_last + pot_last)
        np.testing.assert_array_equal(w2["node_stat"], ns0 + ns1)
        e.output_stat(str(tmp_path), 1)
        rows = np.loadtxt(tmp_path / "heat1")
        assert rows.shape == (m.nnode, 7)
        assert np.any(rows[:, 3] != 0)
    finally:
        e.close()


def test_run_file_writes_the_data_directory(tmp_path, monkeypatch):
    """emc_run_file drives the deck like the reference's main: 8 transient steps
    with a window every 4, then 4 heat steps (Flag_compute_heat) with one window;
    the data/ files are where the plotting scripts look for them.  The heat
    phase counts its steps from 0 again, so its window files append to the
    transient's files of the same number, as in the reference."""
    deck = cfg1_deck(stat_step=STAT_STEP, total_step=8, heat=1, heat_steps=4)
    monkeypatch.chdir(tmp_path)
    (tmp_path / "data").mkdir()
    lib = emcgpu.load_library()
    o = emcgpu.abi.Opts()
    lib.emc_default_opts(C.byref(o))
    table_dir = emcgpu.TABLE_DIR.encode()
    o.table_dir = table_dir
    o.device = 0
    print("[gpu] emc_run_file: 8 steps + 4 heat steps", flush=True)
    assert lib.emc_run_file(deck.encode(), C.byref(o)) == 0
    e = emcgpu.Engine(deck, device=-1)
    m = e.mesh
    e.close()
    d = tmp_path / "data"
    win = emcgpu.read_current(str(d / "current"))
    assert [w["step"] for w in win] == [3, 7, 3]
    assert all(len(w["current"]) == m.ncontact and w["phonon"] > 0 for w in win)
    assert np.loadtxt(d / "pvolume").shape == (m.nnode, 4)
    assert np.loadtxt(d / "pot7").shape == (m.nnode, 4)
    assert np.loadtxt(d / "pot3").shape == (2 * m.nnode, 4)
    assert np.loadtxt(d / "Electron7").shape == (m.nnode, 8)
    assert np.loadtxt(d / "electron_num7").shape == (m.ncell, 4)
    assert np.loadtxt(d / "section_current7").shape == (m.pny, 4)
    heat = np.loadtxt(d / "heat3")
    assert heat.shape == (m.nnode, 7) and np.any(heat[:, 3] != 0)
    assert not (d / "heat7").exists()
    rows = (d / "par_info_restart7").read_text().split("\n")
    assert rows[-2] == "-1 -1 -1" and len(rows) == m.ncell + 2


def test_emc_run_executable_equals_run_file(tmp_path, monkeypatch):
    """`emc_run input.txt` (the process boundary, tables from dirname(exe)/../input)
    writes the same data/ files, byte for byte, as emc_run_file called through the
    library with the package's tables."""
    import filecmp
    import subprocess
    deck = cfg1_deck(stat_step=STAT_STEP, total_step=8, heat=1, heat_steps=4)
    ddir = os.path.dirname(deck)
    a, b = tmp_path / "exe", tmp_path / "lib"
    for d in (a, b):
        d.mkdir()
        for f in ("input.txt", "ldg.txt", "lgrid.txt"):
            (d / f).write_bytes(open(os.path.join(ddir, f), "rb").read())
    exe = os.path.join(REPO, "c-mc_electron_amd", "bin", "emc_run")
    print("[gpu] emc_run executable: 8 steps + 4 heat steps", flush=True)
    r = subprocess.run([exe, "input.txt"], cwd=a, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    monkeypatch.chdir(b)
    (b / "data").mkdir()
    lib = emcgpu.load_library()
    o = emcgpu.abi.Opts()
    lib.emc_default_opts(C.byref(o))
    table_dir = emcgpu.TABLE_DIR.encode()
    o.table_dir = table_dir
    o.device = 0
    assert lib.emc_run_file(b"input.txt", C.byref(o)) == 0
    names = sorted(os.listdir(b / "data"))
    assert "current" in names and "heat3" in names
    assert sorted(os.listdir(a / "data")) == names
    _, mismatch, errors = filecmp.cmpfiles(a / "data", b / "data", names, shallow=False)
    assert not mismatch and not errors, (mismatch, errors)
