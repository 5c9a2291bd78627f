"""get_state with getK as the flight kernels run it (Flight::get_state, compiled
for the host and reached through emc_band_final_state) against the oracle's
restatement of src/ParticleUtils.cpp:48-284 (orc_get_state): the same wave
vector, tetrahedron, number of random numbers and self-scattering flag, bit for
bit, for final energies across both carriers' bands.

The energies include a few meV above the band edges.  There the energy bin's list
is short (8 to 30 entries) and most of its tetrahedra span the energy, but their
DOS share is far below the bin's maximum, so a draw takes hundreds to thousands of
tries (the case behind the flight's rare long launches); the test asserts that such
draws were among those compared."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu


@pytest.fixture(scope="module")
def eng():
    e = emcgpu.Engine(cfg1_deck(), device=-1)
    yield e
    e.close()


CASES = [  # (band, energy eV)
    (0, 0.0002), (0, 0.0009), (0, 0.0016), (0, 0.0031), (0, 0.0055), (0, 0.011), (0, 0.05), (0, 0.4), (0, 2.2),
    (4, 0.0004), (4, 0.0016), (4, 0.0048), (4, 0.03), (4, 0.7), (5, 0.002), (7, 0.003),
]


# serial -1: get_state's loop as written; 0, 1 (the slow flight kernel's value),
# 4, 7, 32: the tries after the first `serial` in chunks of 128 draws, as the
# kernel's wave-parallel form takes them
@pytest.mark.parametrize("serial", [-1, 0, 1, 4, 7, 32])
def test_final_state_equals_the_oracle(eng, serial):
    lib = oracle_lib.load()
    v = eng.band
    rng = np.random.default_rng(11)
    n_per = 60
    most = 0
    for ib, E in CASES:
        keys = rng.integers(1, 2 ** 63, n_per, dtype=np.uint64)
        isym = rng.integers(0, 48, n_per).astype(np.int32)
        energy = np.full(n_per, E)
        iband = np.full(n_per, ib, np.int32)
        k, it, nd, sc = eng.band_final_state(keys, 5, energy, iband, isym, np.zeros(n_per, np.int32), serial)
        ko = np.zeros(3)
        ito = C.c_int(0)
        ndo = C.c_int(0)
        for q in range(n_per):
            ito.value = 0
            r = lib.orc_get_state(C.byref(v), int(keys[q]), 5, float(E), ib, int(isym[q]),
                                  ko.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ito), C.byref(ndo))
            assert (r, ito.value, ndo.value) == (sc[q], it[q], nd[q]), (ib, E, q)
            assert np.array_equal(ko, k[q]), (ib, E, q)
        most = max(most, int(nd.max()))
    # the long loops near the band edges were among them
    assert most > 500
