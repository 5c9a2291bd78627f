"""ctypes binding of oracle/liboracle.so (the CPU restatement used as the checker)."""
import ctypes as C
import os
import subprocess

import numpy as np

from helpers import REPO, emcgpu

abi = emcgpu.abi
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    lib = C.CDLL(ORACLE_LIB)
    P = C.POINTER
    d, i64, u64 = P(C.c_double), P(C.c_int64), P(C.c_uint64)
    sig = {
        "orc_log": (C.c_double, [C.c_double]),
        "orc_sin": (C.c_double, [C.c_double]),
        "orc_cos": (C.c_double, [C.c_double]),
        "orc_philox4x32": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "orc_mix64": (C.c_uint64, [C.c_uint64]),
        "orc_clone_key": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_int]),
        "orc_init_key": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_uniforms": (None, [C.c_uint64, C.c_uint64, C.c_int, d]),
        "orc_fly_step": (C.c_int, [P(abi.BandView), P(abi.MeshView), P(abi.CellFields), P(abi.Particles),
                                   C.c_int64, C.c_int64, C.c_uint64, C.c_int, P(abi.FlyTallies),
                                   i64, i64, i64, i64, i64]),
        "orc_tet_time": (C.c_double, [P(abi.BandView), C.c_int, C.c_int, d, d, P(C.c_int)]),
        "orc_fastsurf": (C.c_double, [P(abi.BandView), C.c_double, C.c_int]),
        "orc_surf": (C.c_double, [P(abi.BandView), C.c_double, C.c_int]),
        "orc_get_state": (C.c_int, [P(abi.BandView), C.c_uint64, C.c_uint64, C.c_double, C.c_int, C.c_int, d,
                                    P(C.c_int), P(C.c_int)]),
        "orc_getisym": (C.c_int, [P(abi.BandView), d, C.c_double, C.c_int, C.c_int, P(C.c_int), P(C.c_int)]),
        "orc_imp_update_k": (C.c_int, [P(abi.BandView), P(abi.MeshView), C.c_uint64, C.c_uint64, d, P(C.c_int),
                                       P(C.c_int), P(C.c_double), C.c_double]),
        "orc_multiple_refresh": (C.c_int64, [P(abi.MeshView), P(abi.Particles), C.c_int64, P(C.c_int32),
                                             P(C.c_int32), C.c_uint64, P(abi.Particles), C.c_int64]),
        "orc_mr_stream_key": (C.c_uint64, [C.c_int64, C.c_int]),
        "orc_mr_child_key": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_int64]),
        "orc_deposit": (C.c_int, [P(abi.BandView), P(abi.MeshView), P(abi.Particles), C.c_int64, i64, i64, i64, i64]),
        "orc_rhs": (None, [P(abi.MeshView), i64, d, d, d]),
        "orc_cell_rho": (None, [P(abi.MeshView), i64, d]),
        "orc_field": (None, [P(abi.MeshView), d, d, d, d]),
        "orc_surface_rates": (None, [P(abi.MeshView), d, d, d, d, d, d, d, d]),
        "orc_pcg": (C.c_int, [P(abi.MeshView), d, d, C.c_double, C.c_int, d]),
        "orc_apply_full": (None, [P(abi.MeshView), d, d]),
        "orc_qc_potential": (C.c_int, [P(abi.MeshView), d, d]),
        "orc_fek": (C.c_double, [P(abi.BandView), C.c_int, C.c_double, C.c_uint64, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def i64ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def fly_step(eng, parts, n, fields, step, flag_heat=0, cap=None, want_heat=False):
    """Run the oracle flight on a copy of `parts` (first n live); returns (new particles, tallies, extras)."""
    lib = load()
    cap = cap or (n + max(4096, n // 8))
    p = parts.copy(cap)
    t = abi.FlyTallies()
    nn, pny = eng.mesh.nnode, eng.mesh.pny
    he = np.zeros(nn, np.int64)
    hh = np.zeros(nn, np.int64)
    rc = np.zeros(pny, np.int64)
    lc = np.zeros(pny, np.int64)
    nout = C.c_int64(0)
    ps = p.struct()
    fs = fields.struct()
    st = lib.orc_fly_step(C.byref(eng.band), C.byref(eng.mesh), C.byref(fs), C.byref(ps), n, cap, step, flag_heat,
                          C.byref(t), i64ptr(he) if want_heat else None, i64ptr(hh) if want_heat else None,
                          i64ptr(rc), i64ptr(lc), C.byref(nout))
    assert st == 0
    p.n = nout.value
    return p, t, {"heat_e": he, "heat_h": hh, "rcur": rc, "lcur": lc}


def deposit(eng, parts, n):
    lib = load()
    nn, nc = eng.mesh.nnode, eng.mesh.ncell
    node_q = np.zeros(nn, np.int64)
    node_stat = np.zeros(10 * nn, np.int64)
    cell_q = np.zeros(3 * nc, np.int64)
    cell_n = np.zeros(nc, np.int64)
    ps = parts.struct()
    lib.orc_deposit(C.byref(eng.band), C.byref(eng.mesh), C.byref(ps), n, i64ptr(node_q), i64ptr(node_stat),
                    i64ptr(cell_q), i64ptr(cell_n))
    return node_q, node_stat, cell_q, cell_n


def multiple_refresh(eng, parts, n, step, desired_e=None, desired_h=None, cap=None):
    """The oracle's MultipleRefresh of `parts` (first n slots): a new Particles."""
    lib = load()
    m = eng.mesh_arrays()
    de = np.ascontiguousarray(m["mr_desired_e"] if desired_e is None else desired_e, np.int32)
    dh = np.ascontiguousarray(m["mr_desired_h"] if desired_h is None else desired_h, np.int32)
    cap = cap if cap is not None else int(n + de.sum() + dh.sum() + 16)
    out = emcgpu.ParticleArrays(cap)
    ps, po = parts.struct(), out.struct()
    k = lib.orc_multiple_refresh(C.byref(eng.mesh), C.byref(ps), n, de.ctypes.data_as(C.POINTER(C.c_int32)),
                                 dh.ctypes.data_as(C.POINTER(C.c_int32)), step, C.byref(po), cap)
    assert k >= 0, "oracle MultipleRefresh: output capacity"
    out.n = int(k)
    return out
