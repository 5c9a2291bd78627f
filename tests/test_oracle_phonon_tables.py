"""Pin the oracle's dispersive phonon final-energy draws (FEKLOAB, FEKLOEM,
FEKLAAB, FEKLAEM, FEKTOAB, FEKTOEM; src/ElectronPhononScatter.cpp:632-1022),
which EPSCAT uses in quantum-correction runs (:148-163).

- The tables are the data files as init_deep reads them (column by column,
  src/ParticleUtils.cpp:750-808), and every column is a probability
  distribution over the final-energy classes.
- Each draw equals a walk written here from the reference's text: the energy
  bin, the class whose running sum first reaches the stream's uniform (pinned
  by the Random123 known answers), and the final energy of that class.
- Over many draws at one energy the classes follow the bin's column.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib
from helpers import PKG, cfg1_deck, emcgpu

TABLES = [("kloem", 201, 120), ("kloab", 201, 110), ("ktoem", 51, 160), ("ktoab", 51, 150),
          ("klaem", 51, 160), ("klaab", 51, 150)]
PROC = {5: ("kloab", "templog"), 6: ("kloem", "templog"), 9: ("klaab", "templaf"), 10: ("klaem", "templaf"),
        11: ("ktoab", "temptof"), 12: ("ktoem", "temptof")}


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(qc=1), device=-1)
    e.tabs = {}
    for t, (name, r, c) in enumerate(TABLES):
        e.tabs[name] = np.ctypeslib.as_array(e.band.ptab[t], shape=(r * c,)).reshape(r, c).copy()
    yield e
    e.close()


def test_tables_as_read(eng):
    for name, r, c in TABLES:
        raw = np.loadtxt(os.path.join(PKG, "input", name + ".txt"))
        np.testing.assert_array_equal(eng.tabs[name], raw.reshape(c, r).T, err_msg=name)
        np.testing.assert_allclose(eng.tabs[name].sum(axis=0), 1.0, atol=2e-6, err_msg=name)


def hw_of(eng, p):
    return getattr(eng.band, PROC[p][1])


def bin_of(p, ie):
    """(ei, bin width) of the reference's piecewise bins."""
    lo, em = p <= 6, p % 2 == 0
    if em:
        rows = ([(3.05, 1.0, 10, 0, 1.0), (5.625, 3.0, 4, 20, 2.0), (13.25, 5.5, 2, 30, 4.0), (28.5, 13.0, 1, 45, 8.0),
                 (None, 28.0, 1, 60, 12.0)] if lo else
                [(3.05, 1.0, 10, 0, 0.25), (5.625, 3.0, 4, 20, 0.5), (13.25, 5.5, 2, 30, 0.5), (None, 13.0, 1, 45, 1.0)])
    else:
        rows = ([(1.05, 0.0, 10, 0, 1.0), (3.625, 1.0, 4, 10, 2.0), (11.25, 3.5, 2, 20, 4.0), (26.5, 11.0, 1, 35, 8.0),
                 (None, 26.0, 1, 50, 12.0)] if lo else
                [(1.05, 0.0, 10, 0, 0.25), (3.625, 1.0, 4, 10, 0.25), (11.25, 3.5, 2, 20, 0.5), (None, 11.0, 1, 35, 1.0)])
    for top, off, scale, base, ed in rows:
        if top is None or ie < top:
            it = int((ie - off) * scale + 0.5) if scale != 1 else int((ie - off) + 0.5)
            it = max(it, 1)
            if top is None:
                if lo:
                    it = min(it, 60)
                elif em:
                    it = min(it, 115)
                elif it > 60:
                    it = 115
            return base + it, ed


def expected(eng, p, eet, ran):
    hw = hw_of(eng, p)
    ie = eet / hw
    lo, em = p <= 6, p % 2 == 0
    if em and ie < 1.0:
        return -1e-10 if p == 10 else 1e-10
    ei, ed = bin_of(p, ie)
    col = eng.tabs[PROC[p][0]][:, ei - 1]
    rows = len(col)
    i, ksum = 0, 0.0
    while ksum < ran:
        i += 1
        if p == 6 and i == rows:
            ksum, i = 1.1, 1
        ksum += col[i - 1]
        if p != 6 and i == rows:
            ksum, i = 1.1, 1
    j, k = i % 2, i // 2
    f = 0.01 if lo else 0.04
    if em:
        e = eet - hw * (1.0 - f * ed * k) if j == 1 else eet - hw * (1.0 + f * ed * k)
    else:
        e = eet + hw * (1.0 + f * ed * k) if j == 1 else eet + hw * (1.0 - f * ed * k)
    if p in (9, 10) and e <= 0.0:
        e = 1.0e-5 * hw
    return e


def uniform(lib, key, step):
    out = np.zeros(1)
    lib.orc_uniforms(key, step, 1, oracle_lib.dptr(out))
    return out[0]


@pytest.mark.parametrize("p", [5, 6, 9, 10, 11, 12])
def test_draws_equal_the_walk(eng, p):
    lib = oracle_lib.load()
    hw = hw_of(eng, p)
    rng = np.random.default_rng(p)
    # energies over every bin of the function, and below one phonon
    for q, ie in enumerate(np.concatenate([rng.uniform(0, 140, 600), [0.2, 0.99, 1.0, 1.04, 3.05, 26.5, 200.0]])):
        eet = ie * hw
        key, step = 1000 * p + q, 7
        got = lib.orc_fek(C.byref(eng.band), p, eet, key, step)
        assert got == expected(eng, p, eet, uniform(lib, key, step)), (p, ie)


@pytest.mark.parametrize("p,ie", [(6, 2.0), (5, 0.5), (9, 5.0), (10, 4.0), (11, 2.0), (12, 20.0)])
def test_classes_follow_the_column(eng, p, ie):
    lib = oracle_lib.load()
    hw = hw_of(eng, p)
    eet = ie * hw
    ei, ed = bin_of(p, ie)
    col = eng.tabs[PROC[p][0]][:, ei - 1]
    n = 20000
    f = 0.01 if p <= 6 else 0.04
    counts = np.zeros(len(col) + 1)
    for key in range(n):
        e = lib.orc_fek(C.byref(eng.band), p, eet, key, 3)
        d = ((eet - e) / hw - 1.0) if p % 2 == 0 else ((e - eet) / hw - 1.0)
        k = int(round(abs(d) / (f * ed)))
        i = 2 * k + (1 if (d <= 0) == (p % 2 == 0) or k == 0 else 0)
        counts[i] += 1
    exp = n * col
    big = exp > 30
    z = (counts[1:][big] - exp[big]) / np.sqrt(exp[big])
    assert np.all(np.abs(z) < 5), (p, np.abs(z).max())
    assert counts[1:][col == 0].sum() <= 2  # only the wrap past the last class lands on class 1
