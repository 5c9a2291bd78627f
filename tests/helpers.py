"""Shared test helpers: import the package, build decks, small populations."""
import os
import sys
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "c-mc_electron_amd")
GOLDEN_DIR = os.path.join(REPO, "tests", "golden")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import emcgpu  # noqa: E402

_DECKS = {}


def deck(name="small", **kw):
    """Path to input.txt of a generated FinFET deck (cached per process)."""
    key = (name, tuple(sorted(kw.items())))
    if key not in _DECKS:
        d = tempfile.mkdtemp(prefix=f"emc_deck_{name}_")
        _DECKS[key] = emcgpu.write_deck(d, **kw)
    return _DECKS[key]


def cfg1_deck(**kw):
    """The 20x40x12-cell (x, y, z) n-FinFET with 20k electrons, Vd 0.5 V, Vg 0.7 V."""
    args = dict(refine=1, electrons=20000, holes=2000, vd=0.5, vg=0.7)
    args.update(kw)
    return deck("cfg1", **args)


def engine(device=-1, **kw):
    return emcgpu.Engine(cfg1_deck(**kw), device=device)


def uniform_fields(eng, ex=0.0, ey=0.0, ez=0.0, rho=None):
    nc = eng.mesh.ncell
    f = emcgpu.CellFieldArrays(nc)
    f.ex[:] = ex
    f.ey[:] = ey
    f.ez[:] = ez
    f.hx[:] = ex
    f.hy[:] = ey
    f.hz[:] = ez
    m = eng.mesh_arrays()
    f.rho[:] = 0.5 * m["da"] if rho is None else rho
    return f


def field_like(eng, seed=0, scale=5.0):
    """A smooth random potential -> fields, Rho and surface rates via the oracle."""
    import oracle_lib
    import ctypes as C
    lib = oracle_lib.load()
    m = eng.mesh
    rng = np.random.default_rng(seed)
    pnx, pny, pnz = m.pnx, m.pny, m.pnz
    j = np.arange(pny)[:, None, None]
    i = np.arange(pnx)[None, :, None]
    k = np.arange(pnz)[None, None, :]
    pot = (scale * (j / pny) + 0.3 * np.sin(2 * np.pi * i / pnx + rng.uniform(0, 6)) +
           0.2 * np.cos(2 * np.pi * k / pnz + rng.uniform(0, 6))).reshape(-1).astype(np.float64)
    f = emcgpu.CellFieldArrays(m.ncell)
    lib.orc_field(C.byref(m), oracle_lib.dptr(pot), oracle_lib.dptr(f.ex), oracle_lib.dptr(f.ey), oracle_lib.dptr(f.ez))
    f.hx[:] = f.ex
    f.hy[:] = f.ey
    f.hz[:] = f.ez
    ma = eng.mesh_arrays()
    f.rho[:] = np.maximum(0.5 * ma["da"], rng.uniform(0, 1, m.ncell) * ma["da"].max() * 0.3 + 1e-6)
    lib.orc_surface_rates(C.byref(m), oracle_lib.dptr(f.ex), oracle_lib.dptr(f.ey), oracle_lib.dptr(f.ez),
                          oracle_lib.dptr(f.rho), oracle_lib.dptr(f.e_rough), oracle_lib.dptr(f.e_sph),
                          oracle_lib.dptr(f.h_rough), oracle_lib.dptr(f.h_sph))
    return pot, f


def sort_by_key(p):
    live = np.nonzero(p.i[:p.n] >= 0)[0]
    order = live[np.argsort(p.key[live], kind="stable")]
    return p.subset(order)


def bulk_deck(cells=2, cell_nm=4, electrons=8000, holes=8000, dt=2e-15, donor=1e21):
    """input.txt of a bulk silicon box: cells^3 cells of cell_nm nm, every outer
    face PERIOD (a carrier leaving a boundary cell re-enters it on the opposite
    face, CellHit.cpp Period()), donors 1e21 m^-3 (below the Brooks-Herring
    threshold, so no impurity scattering), surface scattering off.  The one
    contact only satisfies the grammar: these decks never solve Poisson."""
    key = ("bulk", cells, cell_nm, electrons, holes, dt, donor)
    if key in _DECKS:
        return _DECKS[key]
    d = tempfile.mkdtemp(prefix="emc_deck_bulk_")
    L = cells * cell_nm
    with open(os.path.join(d, "lgrid.txt"), "w") as f:
        for _ in range(3):
            f.write("%d\n" % (cells + 1))
            for c in range(cells + 1):
                f.write("%.10g\n" % (c * cell_nm * 1e-3))
    ldg = f"""default_par_number 1 1
region 0 {L} 0 {L} 0 {L} SILICON
donor 0 {L} 0 {L} 0 {L} {donor:g}
acceptor 0 {L} 0 {L} 0 {L} 0
motioncube 0 {L} 0 {L} 0 {L} PASS PASS PASS PASS PASS PASS
motionplane 0 0 0 {L} 0 {L} UP PERIOD
motionplane {L} {L} 0 {L} 0 {L} DOWN PERIOD
motionplane 0 {L} 0 0 0 {L} LEFT PERIOD
motionplane 0 {L} {L} {L} 0 {L} RIGHT PERIOD
motionplane 0 {L} 0 {L} 0 0 FRONT PERIOD
motionplane 0 {L} 0 {L} {L} {L} BACK PERIOD
contact
  1 0
  0 {L} 0 0 0 {L}
  0
end
"""
    with open(os.path.join(d, "ldg.txt"), "w") as f:
        f.write(ldg)
    with open(os.path.join(d, "input.txt"), "w") as f:
        f.write(f"""gridFile=lgrid.txt
device_file_name=ldg.txt
total_step=1000
ElectronNumber={electrons}
HoleNumber={holes}
Temperature=300
dt={dt:g}
stat_step=1000
stat_heat_step=1000
heat_steps=0
Flag_QuantumCorrection=0
Flag_MultipleRefresh=0
Flag_compute_heat=0
Flag_SurfaceScatter=0
mr_step=10
""")
    _DECKS[key] = os.path.join(d, "input.txt")
    return _DECKS[key]
