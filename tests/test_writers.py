"""The data/ writers against the reference's formats (CPU, no GPU).

output_stat / current_scatter_info (src/mcmodel.cpp:971-1161, 756-842),
output_for_restart (src/mcmodel.cpp:895-965) and print_p_data
(src/utils.cpp:311-338) print doubles through an ofstream in its default
format; the first test pins that format to printf's %g with g++ itself, so
the expected lines below can be written with Python's "%g".  The restart
writer reproduces rows of a file the reference wrote (tests/golden), and the
current file is read back the way finfet/scripts/current.py and
plot_current.py read it.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from helpers import cfg1_deck, emcgpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAT_STEP = 4


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(stat_step=STAT_STEP), device=-1)
    yield e
    e.close()


def g(v):
    return "%g" % v


def test_ostream_default_format_is_printf_g(tmp_path):
    """std::ostream << double with default flags prints what printf("%g") prints."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    vals = [0.0, -0.0, 1.0, 0.1, 1e-5, 1.5e-5, 123456.0, 1234567.0, 999999.5, 0.000123456789, -2.5e-300,
            6.02214076e23, 1.0 / 3.0, 100000.0, 1e6, 2.0 ** 52, float("inf")]
    src = tmp_path / "p.cpp"
    src.write_text("#include <iostream>\n#include <cstdlib>\nint main(int c, char** v) {\n"
                   "  for (int i = 1; i < c; ++i) std::cout << std::strtod(v[i], nullptr) << '\\n';\n}\n")
    exe = tmp_path / "p"
    subprocess.run([gxx, "-O1", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)] + [repr(v) for v in vals], check=True, capture_output=True, text=True).stdout
    assert out.split("\n")[:-1] == [g(v) for v in vals]


def node_order(m):
    """print_p_data's loop order: i, then k, then j fastest."""
    return [(i, j, k) for i in range(m.pnx) for k in range(m.pnz) for j in range(m.pny)]


def node_index(m, i, j, k):
    return (j * m.pnx + i) * m.pnz + k


def test_pvolume_file(eng, tmp_path):
    eng.output_pvolume(str(tmp_path))
    m = eng.mesh
    pvol = eng.mesh_arrays()["pvol"]
    lines = (tmp_path / "pvolume").read_text().split("\n")
    assert lines[-1] == ""
    want = [f"{i} {j} {k} {g(pvol[node_index(m, i, j, k)] * m.conc0)}" for i, j, k in node_order(m)]
    assert lines[:-1] == want
    # the plotting scripts read it with numpy.loadtxt
    assert np.loadtxt(tmp_path / "pvolume").shape == (m.nnode, 4)


def known_window(eng, seed=0):
    """A window whose every printed value follows from the formulas in a line."""
    m = eng.mesh
    nn, nc = m.nnode, m.ncell
    rng = np.random.default_rng(seed)
    w = eng.new_window()
    w["stat_pot"][:] = rng.uniform(-1, 1, nn) * STAT_STEP
    w["stat_qc_pot"][:] = rng.uniform(-1, 1, nn) * STAT_STEP
    ns = w["node_stat"].reshape(10, nn)
    q = rng.integers(1, 1 << 40, (10, nn))
    q[[0, 1, 2, 5, 6, 7]] -= 1 << 39                      # signed velocity sums
    q[3, ::5] = 0                                          # |q| below MY_ZERO: zero row
    q[8, 1::7] = 0
    ns[:] = q
    w["cell_count"][:] = rng.integers(0, 400, nc)
    w["heat_e"][:] = rng.integers(0, 1 << 44, nn)
    w["heat_h"][:] = rng.integers(0, 1 << 44, nn)
    w["rcur"][:] = rng.integers(-(1 << 42), 1 << 42, m.pny)
    w["lcur"][:] = rng.integers(-(1 << 42), 1 << 42, m.pny)
    t = emcgpu.abi.FlyTallies()
    for c in range(m.ncontact):
        t.contact[c][0] = 100 + c                          # NumParGen
        t.contact[c][1] = int(rng.integers(0, 1 << 45))    # CharGen
        t.contact[c][2] = int(rng.integers(0, 1 << 45))    # EnergyGen
        t.contact[c][3] = 200 + c                          # NumParCatch
        t.contact[c][4] = int(rng.integers(0, 1 << 45))    # CharCatch
    t.n_phonon = 12345
    t.n_impurity = 678
    w["tallies"] = t
    return w


def expected_current(eng, w, step):
    m = eng.mesh
    t = w["tallies"]
    dts = m.dt * STAT_STEP
    fx = emcgpu.FX_CONTACT
    out = [f"step = {step}"]
    for c in range(m.ncontact):
        gen, cat, egen = t.contact[c][1] / fx, t.contact[c][4] / fx, t.contact[c][2] / fx
        out += [f"icont = {c}",
                f"CharGen = {g(gen)} NumParGen = {t.contact[c][0]} CharCatch = {g(cat)} NumParCatch = {t.contact[c][3]}",
                f"EnergyGen= {g(egen)}", "EnergyCatch= 0", f"EnergyCurrent= {g((0.0 - egen) / dts * m.pot0)}",
                f"Current = {g((gen - cat) / dts * m.curr0 * 1e-2)}"]
    out += [f"times of phonon scatter : {t.n_phonon}", f"times of impurity scatter : {t.n_impurity}"]
    return out


def expected_carrier(eng, w, t):
    """Electron (t=0) / Hole (t=1) rows and the densities output_stat keeps."""
    m = eng.mesh
    nn = m.nnode
    ns = w["node_stat"].reshape(10, nn) / emcgpu.FX_STAT
    pvol = eng.mesh_arrays()["pvol"]
    rows, dens = [], np.zeros(nn)
    for i, j, k in node_order(m):
        n = node_index(m, i, j, k)
        q = ns[5 * t + 3, n]
        a = abs(q)
        v = [0.0] * 4
        if a > 1e-13:
            v = [ns[5 * t + 0, n] / a * m.velo0, ns[5 * t + 1, n] / a * m.velo0, ns[5 * t + 2, n] / a * m.velo0,
                 ns[5 * t + 4, n] / a * m.pot0]
        d = q / pvol[n] / STAT_STEP * m.conc0 if pvol[n] > 1e-13 else 0.0
        dens[n] = d
        rows.append(f"{i} {j} {k} " + " ".join(g(x) for x in v + [d]))
    return rows, dens


def read_lines(path):
    s = path.read_text()
    assert s.endswith("\n")
    return s.split("\n")[:-1]


def test_window_files_from_known_window(eng, tmp_path):
    m = eng.mesh
    w = known_window(eng)
    eng.write_stat(w, str(tmp_path), 7)
    assert read_lines(tmp_path / "current") == expected_current(eng, w, 7)
    dts = m.dt * STAT_STEP
    fx = emcgpu.FX_CONTACT
    want = []
    for y in range(m.pny):
        r, lc = w["rcur"][y] / fx, w["lcur"][y] / fx
        want.append(f"{y} {g(r / dts * m.curr0)} {g(lc / dts * m.curr0)} {g((r - lc) / dts * m.curr0)}")
    assert read_lines(tmp_path / "section_current7") == want
    want = [f"{i} {j} {k} {g(w['cell_count'][(j * m.cnx + i) * m.cnz + k] / STAT_STEP)}"
            for i in range(m.cnx) for j in range(m.cny) for k in range(m.cnz)]
    assert read_lines(tmp_path / "electron_num7") == want
    want = [f"{i} {j} {k} {g(w['stat_pot'][node_index(m, i, j, k)] / STAT_STEP * m.pot0)}" for i, j, k in node_order(m)]
    assert read_lines(tmp_path / "pot7") == want
    want = [f"{i} {j} {k} {g(w['stat_qc_pot'][node_index(m, i, j, k)] / STAT_STEP * m.pot0)}"
            for i, j, k in node_order(m)]
    assert read_lines(tmp_path / "qc_pot7") == want
    # without a corrected-potential sum the qc_pot file holds the Poisson one
    (tmp_path / "noqc").mkdir()
    eng.write_stat({"stat_pot": w["stat_pot"]}, str(tmp_path / "noqc"), 7)
    assert read_lines(tmp_path / "noqc" / "qc_pot7") == read_lines(tmp_path / "pot7")
    e_rows, e_dens = expected_carrier(eng, w, 0)
    h_rows, h_dens = expected_carrier(eng, w, 1)
    assert read_lines(tmp_path / "Electron7") == e_rows
    assert read_lines(tmp_path / "Hole7") == h_rows
    assert not (tmp_path / "heat7").exists(), "heat file written outside the heat phase"

    # a second window: data/current is appended, and the heat file uses the
    # densities kept from the previous window (stat_chargeE_saved)
    w2 = known_window(eng, seed=1)
    eng.write_stat(w2, str(tmp_path), 11, heat_file=True)
    assert read_lines(tmp_path / "current") == expected_current(eng, w, 7) + expected_current(eng, w2, 11)
    ns = w2["node_stat"].reshape(10, m.nnode) / emcgpu.FX_STAT
    want = []
    for i, j, k in node_order(m):
        n = node_index(m, i, j, k)
        qe, qh = ns[3, n], ns[8, n]
        v1 = w2["heat_e"][n] / emcgpu.FX_HEAT * m.pot0 / qe if abs(qe) > 1e-13 else 0.0
        v2 = w2["heat_h"][n] / emcgpu.FX_HEAT * m.pot0 / qh if abs(qh) > 1e-13 else 0.0
        want.append(f"{i} {j} {k} {g(v1)} {g(v1 * e_dens[n])} {g(v2)} {g(v2 * h_dens[n])}")
    assert read_lines(tmp_path / "heat11") == want
    # a window written at step 0 starts data/current afresh
    eng.write_stat({"tallies": w["tallies"]}, str(tmp_path), 0)
    assert read_lines(tmp_path / "current") == expected_current(eng, w, 0)


def test_current_file_reads_like_the_scripts(eng, tmp_path):
    """Five windows of data/current parsed as current.py and plot_current.py parse
    them; current_estimate's mean is current.py's estimate."""
    m = eng.mesh
    assert m.ncontact >= 2
    cur = []
    for s in range(5):
        w = known_window(eng, seed=10 + s)
        eng.write_stat({"tallies": w["tallies"]}, str(tmp_path), s * STAT_STEP + STAT_STEP - 1)
        t = w["tallies"]
        cur.append([(t.contact[c][1] - t.contact[c][4]) / emcgpu.FX_CONTACT / (m.dt * STAT_STEP) * m.curr0 * 1e-2
                    for c in range(m.ncontact)])
    # current.py: windows counted by "step" lines, icont = split[2], Current = split[2]
    steps, total, icont, items = 0, 0.0, 0, 1
    for line in (tmp_path / "current").read_text().split("\n"):
        if re.match("step", line):
            steps += 1
        if re.match("icont", line):
            icont = int(re.split(r"\s+", line)[2])
        if icont <= 1 and steps > items and re.match("Current", line):
            total += abs(float(re.split(r"\s+", line)[2]))
    script = total / (2 * (steps - items))
    est = emcgpu.current_estimate(str(tmp_path / "current"), contacts=(0, 1), skip=1)
    assert est["windows"] == 4
    assert est["mean"] == pytest.approx(script, rel=1e-12)
    x = np.array([(abs(float(g(c[0]))) + abs(float(g(c[1])))) / 2 for c in cur[1:]])
    assert est["mean"] == pytest.approx(x.mean(), rel=1e-12)
    assert est["half_width"] == pytest.approx(3.182446305284263 * x.std(ddof=1) / 2.0, rel=1e-9)  # t(0.975, 3)
    win = emcgpu.read_current(str(tmp_path / "current"))
    assert [w["step"] for w in win] == [3, 7, 11, 15, 19]
    assert win[2]["phonon"] == 12345 and win[2]["impurity"] == 678


def parse_rows(path):
    rows = {}
    for line in open(path):
        v = line.split()
        if len(v) == 7:
            rows[tuple(int(x) for x in v[:3])] = line.rstrip("\n")
    return rows


def test_restart_file_reproduces_reference_rows(eng, tmp_path):
    """Particles placed to give each cell the charges and counts of a reference
    restart file: the writer prints those rows character for character, in the
    reference's cell order, then the -1 -1 -1 terminator."""
    m = eng.mesh
    ref = parse_rows(os.path.join(GOLDEN, "par_info_restart_slice.txt"))
    cells = [c for c in ref if c[0] < m.cnx and c[1] < m.cny and c[2] < m.cnz]
    assert len(cells) == 2400
    n = sum(int(ref[c].split()[4]) + int(ref[c].split()[6]) for c in cells)
    p = emcgpu.ParticleArrays(n)
    q = 0
    for c in cells:
        v = ref[c].split()
        for t, (qc, num) in enumerate(((float(v[3]), int(v[4])), (float(v[5]), int(v[6])))):
            for r in range(num):
                p.i[q], p.j[q], p.k[q] = c
                p.ptype[q] = t
                p.charge[q] = (-qc if t == 0 else qc) / num
                p.key[q] = q * 2654435761 % (1 << 61)
                q += 1
    assert q == n
    eng.write_restart(p, n, str(tmp_path), 99999)
    lines = read_lines(tmp_path / "par_info_restart99999")
    assert lines[-1] == "-1 -1 -1"
    order = [(i, j, k) for i in range(m.cnx) for j in range(m.cny) for k in range(m.cnz)]
    assert len(lines) == len(order) + 1
    for line, c in zip(lines, order):
        assert line == ref.get(c, f"{c[0]} {c[1]} {c[2]} 0 0 0 0"), c
    # empty slots are skipped; a particle outside the mesh is refused
    p.i[0] = -1
    eng.write_restart(p, n, str(tmp_path), 1)
    p.i[0] = m.cnx
    with pytest.raises(emcgpu.EmcError):
        eng.write_restart(p, n, str(tmp_path), 2)


def test_writers_fail_loudly_on_a_missing_directory(eng, tmp_path):
    with pytest.raises(emcgpu.EmcError) as e:
        eng.output_pvolume(str(tmp_path / "no_such_dir"))
    assert e.value.code == emcgpu.abi.EMC_ERR_IO


def test_device_entry_points_need_a_gpu(eng, tmp_path):
    for call in (lambda: eng.output_stat(str(tmp_path), 0), lambda: eng.output_for_restart(str(tmp_path), 0),
                 lambda: eng.heat_begin(), lambda: eng.heat_step(1), lambda: eng.stat_window()):
        with pytest.raises(emcgpu.EmcError):
            call()


def restart_deck(tmp_path, restart_file, step=8):
    """The 20x40x12 deck with Flag_restart pointing at restart_file."""
    src = os.path.dirname(cfg1_deck(stat_step=STAT_STEP))
    d = tmp_path / "deck"
    shutil.copytree(src, d)
    with open(d / "input.txt", "a") as f:
        f.write(f"Flag_restart=1\nrestart_step={step}\nrestart_filename={restart_file}\n")
    return str(d / "input.txt")


def cell_groups(p, m):
    """{(cell, type): (count, charges)} of a host population."""
    live = np.nonzero(p.i[:p.n] >= 0)[0]
    c = (p.j[live].astype(np.int64) * m.cnx + p.i[live]) * m.cnz + p.k[live]
    out = {}
    for s, cc in zip(live, c):
        out.setdefault((int(cc), int(p.ptype[s])), []).append(p.charge[s])
    return out


def test_restart_import_of_a_reference_file(eng, tmp_path):
    """Flag_restart with a file the reference wrote (read_par_info_for_restart,
    src/particle_init.cpp:128-170): cells listed in the file get its counts; away
    from contacts each particle carries the file's charge / count (electrons
    negative), at contacts the neutral charge / count; cells not in the file keep
    the computed counts, with zero charge away from contacts."""
    m = eng.mesh
    ref = parse_rows(os.path.join(GOLDEN, "par_info_restart_slice.txt"))
    eng.particles_init()
    base = cell_groups(eng.download(), m)
    r = emcgpu.Engine(restart_deck(tmp_path, os.path.join(GOLDEN, "par_info_restart_slice.txt")), device=-1)
    try:
        r.particles_init()
        got = cell_groups(r.download(), m)
    finally:
        r.close()
    ma = eng.mesh_arrays()
    silicon = {cc for cc, _ in base}
    listed = {(c[1] * m.cnx + c[0]) * m.cnz + c[2]: ref[c].split() for c in ref}
    seen_restart = seen_contact = 0
    for cc in silicon:
        for t in (0, 1):
            qb = base[(cc, t)]
            if cc not in listed:
                # computed count; the charge of a cell the file does not list is 0
                # unless the cell is at a contact (the reference's charge vector
                # starts zeroed)
                assert len(got[(cc, t)]) == len(qb)
                assert got[(cc, t)] == (qb if ma["attached"][cc] > 0 else [0.0] * len(qb))
                continue
            v = listed[cc]
            num = int(v[4 + 2 * t])
            if num == 0:
                assert (cc, t) not in got
                continue
            q = got[(cc, t)]
            assert len(q) == num
            if ma["attached"][cc] > 0:
                want = qb[0] * len(qb) / num
                np.testing.assert_allclose(q, want, rtol=1e-14)
                seen_contact += 1
            else:
                want = (-float(v[3]) if t == 0 else float(v[5])) / num
                assert all(x == want for x in q)
                seen_restart += 1
    assert seen_restart > 100 and seen_contact > 0
    # a missing file and a file without its terminator are errors
    for bad in (str(tmp_path / "missing"), None):
        if bad is None:
            bad = str(tmp_path / "cut")
            with open(bad, "w") as f:
                f.write("0 0 0 1 1 1 1\n")
        r = emcgpu.Engine(restart_deck(tmp_path / ("d" + os.path.basename(bad)), bad), device=-1)
        try:
            with pytest.raises(emcgpu.EmcError) as e:
                r.particles_init()
            assert e.value.code == emcgpu.abi.EMC_ERR_IO
        finally:
            r.close()
