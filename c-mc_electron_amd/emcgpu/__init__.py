"""Python host wrapper over libemc.so (the C ABI in include/emc.h).

The engine runs the flight, deposit, field and Poisson kernels on the GPU.
There is no CPU fallback: without the built library or a GPU the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (FX_CONTACT, FX_HEAT, FX_Q, FX_STAT, CELL_FIELDS, PARTICLE_F64, PARTICLE_I32)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libemc.so")
TABLE_DIR = os.path.join(PKG_DIR, "input")
DECK_DIR = os.path.join(PKG_DIR, "decks")

_lib = None


class EmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"emc error {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise EmcError(abi.EMC_ERR_NOGPU, f"{path} is not built; run __graft_entry__.build()")
        lib = C.CDLL(path)
        abi.declare(lib)
        _lib = lib
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _view(ptr, n, dtype):
    if not ptr or n <= 0:
        return np.zeros(0, dtype=dtype)
    ct = np.ctypeslib.as_ctypes_type(dtype)
    buf = C.cast(ptr, C.POINTER(ct * n)).contents
    return np.frombuffer(buf, dtype=dtype, count=n)


class ParticleArrays:
    """Host SoA buffers matching emc_particles."""

    def __init__(self, n):
        self.n = n
        for f in PARTICLE_F64:
            setattr(self, f, np.zeros(n, np.float64))
        for f in PARTICLE_I32:
            setattr(self, f, np.zeros(n, np.int32))
        self.key = np.zeros(n, np.uint64)

    def struct(self):
        s = abi.Particles()
        for f in PARTICLE_F64:
            setattr(s, f, _ptr(getattr(self, f), C.c_double))
        for f in PARTICLE_I32:
            setattr(s, f, _ptr(getattr(self, f), C.c_int32))
        s.key = _ptr(self.key, C.c_uint64)
        return s

    def subset(self, idx):
        out = ParticleArrays(len(idx))
        for f in PARTICLE_F64 + PARTICLE_I32 + ("key",):
            getattr(out, f)[:] = getattr(self, f)[idx]
        return out

    def copy(self, cap=None):
        cap = self.n if cap is None else cap
        out = ParticleArrays(cap)
        for f in PARTICLE_F64 + PARTICLE_I32 + ("key",):
            getattr(out, f)[:self.n] = getattr(self, f)[:self.n]
        out.n = self.n
        return out

    def live(self):
        return self.subset(np.nonzero(self.i[:self.n] >= 0)[0])


class CellFieldArrays:
    def __init__(self, ncell):
        for f in CELL_FIELDS:
            setattr(self, f, np.zeros(ncell, np.float64))

    def struct(self):
        s = abi.CellFields()
        for f in CELL_FIELDS:
            setattr(s, f, _ptr(getattr(self, f), C.c_double))
        return s


class Engine:
    """One device description (decks) plus its GPU state."""

    def __init__(self, input_txt, device=0, seed=1, table_dir=TABLE_DIR, capacity_factor=1.25,
                 poisson_rtol=1e-10, poisson_maxit=20000, poisson_fp32=None, poisson_inner_rtol=1e-5):
        self.lib = load_library()
        o = abi.Opts()
        self.lib.emc_default_opts(C.byref(o))
        self._table_dir = table_dir.encode()
        o.table_dir = self._table_dir
        o.device = device
        o.seed = seed
        o.capacity_factor = capacity_factor
        o.poisson_rtol = poisson_rtol
        o.poisson_maxit = poisson_maxit
        if poisson_fp32 is not None:
            # 0 / False: FP64 throughout; 1 / True: FP32 bulk with FP64 refinement;
            # 2: FP64 CG with the FP32 cycle (the library's default)
            o.poisson_fp32 = int(poisson_fp32)
        o.poisson_inner_rtol = poisson_inner_rtol
        st = C.c_int(0)
        self.ctx = self.lib.emc_create(os.fsencode(input_txt), C.byref(o), C.byref(st))
        if st.value != 0:
            msg = self.lib.emc_last_error(self.ctx).decode()
            self.lib.emc_destroy(self.ctx)
            self.ctx = None
            raise EmcError(st.value, msg)
        self.band = abi.BandView()
        self._check(self.lib.emc_band_view_get(self.ctx, C.byref(self.band)))
        self.mesh = abi.MeshView()
        self._check(self.lib.emc_mesh_view_get(self.ctx, C.byref(self.mesh)))

    def close(self):
        if self.ctx:
            self.lib.emc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise EmcError(st, self.lib.emc_last_error(self.ctx).decode())

    # ---- views
    def band_arrays(self):
        b = self.band
        nt, nk = b.nt, b.nk
        v = {}
        for f, n in (("xkk", nk), ("ykk", nk), ("zkk", nk), ("eek", nk), ("vgt", nt), ("gamtet", nt),
                     ("faclow", nt), ("facmedium", nt), ("fachigh", nt), ("maxaovtet", nt),
                     ("xkct", nt), ("ykct", nt), ("zkct", nt), ("vt", 4 * nt), ("dnt", 16 * nt),
                     ("energy", 7501), ("dos", 8 * 7501), ("sumdos", 7501 * 3), ("dostetmax", 8 * 7501),
                     ("sumscatt", 7501 * 8), ("scatte", 14 * 16), ("dose", 14 * 4 * 7501),
                     ("scattiie", 16 * 7501), ("scatth", 4 * 9), ("dosh", 4 * 3 * 7501), ("scattiih", 9 * 7501),
                     ("xbz", 240), ("ybz", 240), ("zbz", 240), ("tsphere", 4 * nt)):
            v[f] = _view(getattr(b, f), n, np.float64)
        for f, n in (("tet", 4 * nt), ("neib", 4 * nt), ("ibt", nt), ("matsym", 288), ("indmat", 729),
                     ("newisym", 240), ("ntlist", 325 * 8), ("ptlist", 325 * 8), ("tlist", b.tlist_len),
                     ("nslist", 325 * 8), ("pslist", 325 * 8), ("slist", b.slist_len),
                     ("ntlists", 60 * 8), ("ptlists", 60 * 8), ("tlists", b.tlists_len)):
            v[f] = _view(getattr(b, f), n, np.int32)
        return v

    def mesh_arrays(self):
        m = self.mesh
        nc, nn = m.ncell, m.nnode
        v = {}
        for f, n in (("lx", m.pnx), ("ly", m.pny), ("lz", m.pnz), ("dx", m.cnx), ("dy", m.cny), ("dz", m.cnz),
                     ("cvol", nc), ("da", nc), ("pvol", nn), ("pdop", nn), ("vadd", nn), ("dirichlet", nn),
                     ("diag", nn), ("cx", nn), ("cy", nn), ("cz", nn), ("rhs_corr", nn), ("stencil", 7 * nn)):
            v[f] = _view(getattr(m, f), n, np.float64)
        for f, n in (("material", nc), ("attached", nc), ("nearest_surf", nc), ("eeff_dir", nc),
                     ("surface_dir", max(1, m.nsurf)), ("mr_desired_e", nc), ("mr_desired_h", nc)):
            v[f] = _view(getattr(m, f), n, np.int32)
        for f, n in (("motion", 6 * nc), ("insurf", nc), ("node_kind", nn), ("qcregion", nc)):
            v[f] = _view(getattr(m, f), n, np.uint8)
        return v

    # ---- particles
    def particles_init(self):
        n = C.c_int64(0)
        self._check(self.lib.emc_particles_init(self.ctx, C.byref(n)))
        return n.value

    def count(self):
        n = C.c_int64(0)
        self._check(self.lib.emc_particles_count(self.ctx, C.byref(n)))
        return n.value

    def upload(self, p, n=None):
        n = p.n if n is None else n
        s = p.struct()
        self._check(self.lib.emc_particles_upload(self.ctx, C.byref(s), n))

    def download(self):
        n = self.count()
        p = ParticleArrays(n)
        s = p.struct()
        got = C.c_int64(0)
        self._check(self.lib.emc_particles_download(self.ctx, C.byref(s), n, C.byref(got)))
        p.n = got.value
        return p

    # ---- fields
    def fields_set(self, f):
        s = f.struct()
        self._check(self.lib.emc_fields_set(self.ctx, C.byref(s)))

    def fields_get(self):
        f = CellFieldArrays(self.mesh.ncell)
        s = f.struct()
        self._check(self.lib.emc_fields_get(self.ctx, C.byref(s)))
        return f

    # ---- kernels
    def fly_step(self, step, flag_heat=0):
        t = abi.FlyTallies()
        self._check(self.lib.emc_fly_step(self.ctx, step, flag_heat, C.byref(t)))
        return t

    def deposit(self):
        nn, nc = self.mesh.nnode, self.mesh.ncell
        node_q = np.zeros(nn, np.int64)
        node_stat = np.zeros(10 * nn, np.int64)
        cell_q = np.zeros(3 * nc, np.int64)
        cell_n = np.zeros(nc, np.int64)
        self._check(self.lib.emc_deposit(self.ctx, _ptr(node_q, C.c_int64), _ptr(node_stat, C.c_int64),
                                         _ptr(cell_q, C.c_int64), _ptr(cell_n, C.c_int64)))
        return node_q, node_stat, cell_q, cell_n

    def rhs_from_charge(self, node_q):
        node_q = np.ascontiguousarray(node_q, np.int64)
        rhs = np.zeros(self.mesh.nnode)
        self._check(self.lib.emc_rhs_from_charge(self.ctx, _ptr(node_q, C.c_int64), _ptr(rhs, C.c_double)))
        return rhs

    def poisson_solve(self, rhs, x0=None, rtol=1e-10, maxit=20000):
        rhs = np.ascontiguousarray(rhs, np.float64)
        sol = np.zeros(self.mesh.nnode) if x0 is None else np.array(x0, np.float64)
        it = C.c_int(0)
        rr = C.c_double(0)
        self._check(self.lib.emc_poisson_solve(self.ctx, _ptr(rhs, C.c_double), _ptr(sol, C.c_double), rtol, maxit,
                                               C.byref(it), C.byref(rr)))
        return sol, it.value, rr.value

    def poisson_solve_mixed(self, rhs, x0=None, rtol=1e-10, maxit_outer=20, inner_rtol=1e-5):
        """The solve with an FP32 bulk and FP64 iterative refinement: (sol, outer
        steps, FP32 CG iterations, relative residual)."""
        rhs = np.ascontiguousarray(rhs, np.float64)
        sol = np.zeros(self.mesh.nnode) if x0 is None else np.array(x0, np.float64)
        o, i = C.c_int(0), C.c_int(0)
        rr = C.c_double(0)
        self._check(self.lib.emc_poisson_solve_mixed(self.ctx, _ptr(rhs, C.c_double), _ptr(sol, C.c_double), rtol,
                                                     maxit_outer, inner_rtol, C.byref(o), C.byref(i), C.byref(rr)))
        return sol, o.value, i.value, rr.value

    def poisson_solve_noalloc(self, rhs, x0=None, rtol=1e-10, maxit=20000):
        """poisson_solve through emc_poisson_solve_noalloc (work vectors in the
        library's static device array; nothing allocated per call)."""
        rhs = np.ascontiguousarray(rhs, np.float64)
        sol = np.zeros(self.mesh.nnode) if x0 is None else np.array(x0, np.float64)
        it = C.c_int(0)
        rr = C.c_double(0)
        self._check(self.lib.emc_poisson_solve_noalloc(self.ctx, _ptr(rhs, C.c_double), _ptr(sol, C.c_double), rtol,
                                                       maxit, C.byref(it), C.byref(rr)))
        return sol, it.value, rr.value

    def tet_index(self):
        """The point-location octree: (roots, nodes, entries) numpy copies, or None."""
        b = self.band
        if b.tindex_nodes <= 0:
            return None
        rdt = np.dtype([("lo", "<f8", 3), ("hi", "<f8", 3), ("node", "<i4"), ("pad", "<i4")])
        edt = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("t", "<i4"), ("bins", "<u4")])
        roots = np.frombuffer((C.c_char * (8 * rdt.itemsize)).from_address(b.tindex_root), rdt).copy()
        nodes = _view(b.tindex_node, 2 * b.tindex_nodes, np.int32).reshape(-1, 2).copy()
        ent = np.frombuffer((C.c_char * (b.tindex_entries * edt.itemsize)).from_address(b.tindex_entry), edt).copy()
        return roots, nodes, ent

    def checkpoint_save(self, path):
        self._check(self.lib.emc_checkpoint_save(self.ctx, os.fsencode(path)))

    def checkpoint_load(self, path):
        self._check(self.lib.emc_checkpoint_load(self.ctx, os.fsencode(path)))

    def multiple_refresh(self, step):
        """MultipleRefresh of the device population; returns (live before, after)."""
        b, a = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.emc_multiple_refresh(self.ctx, step, C.byref(b), C.byref(a)))
        return b.value, a.value

    def band_locate(self, k, energy, iband, itet_hint):
        """Getisym on the host tables for points k [n, 3]: returns (itet, isym),
        itet -1 where no tetrahedron holds the point."""
        k = np.ascontiguousarray(k, np.float64).reshape(-1, 3)
        n = len(k)
        e = np.ascontiguousarray(energy, np.float64).reshape(n)
        ib = np.ascontiguousarray(iband, np.int32).reshape(n)
        it = np.array(itet_hint, np.int32).reshape(n).copy()
        sy = np.zeros(n, np.int32)
        self._check(self.lib.emc_band_locate(self.ctx, n, _ptr(k, C.c_double), _ptr(e, C.c_double), _ptr(ib, C.c_int),
                                             _ptr(it, C.c_int), _ptr(sy, C.c_int)))
        return it, sy

    def band_final_state(self, key, step, energy, iband, isym, itet, serial=-1):
        """get_state with getK on the host tables for n draws: returns (k [n, 3],
        itet, ndraws, selfsc).  serial >= 0: the tries after the first `serial` in
        chunks of 128 draws, as the slow flight kernel takes them."""
        key = np.ascontiguousarray(key, np.uint64).reshape(-1)
        n = len(key)
        e = np.ascontiguousarray(energy, np.float64).reshape(n)
        ib = np.ascontiguousarray(iband, np.int32).reshape(n)
        sy = np.ascontiguousarray(isym, np.int32).reshape(n)
        it = np.array(itet, np.int32).reshape(n).copy()
        k = np.zeros((n, 3))
        nd = np.zeros(n, np.int32)
        sc = np.zeros(n, np.int32)
        self._check(self.lib.emc_band_final_state(self.ctx, n, _ptr(key, C.c_uint64), step, _ptr(e, C.c_double),
                                                  _ptr(ib, C.c_int), _ptr(sy, C.c_int), _ptr(k, C.c_double),
                                                  _ptr(it, C.c_int), _ptr(nd, C.c_int), _ptr(sc, C.c_int),
                                                  int(serial)))
        return k, it, nd, sc

    def mg_levels(self):
        """The Poisson preconditioner's multigrid hierarchy: list of (dims, a27 or
        None, omega) per level; a27 has shape (27, n)."""
        out = []
        level = 0
        while True:
            dims = (C.c_int * 3)()
            om = C.c_double(0)
            st = self.lib.emc_poisson_mg_level(self.ctx, level, dims, None, C.byref(om))
            if st != 0:
                break
            a27 = None
            if level > 0:
                a27 = np.zeros(27 * dims[0] * dims[1] * dims[2])
                self._check(self.lib.emc_poisson_mg_level(self.ctx, level, dims, _ptr(a27, C.c_double), None))
                a27 = a27.reshape(27, -1)
            out.append((tuple(dims), a27, om.value))
            level += 1
        return out

    def compute_field(self, pot, rho):
        pot = np.ascontiguousarray(pot, np.float64)
        rho = np.ascontiguousarray(rho, np.float64)
        self._check(self.lib.emc_compute_field(self.ctx, _ptr(pot, C.c_double), _ptr(rho, C.c_double)))
        return self.fields_get()

    def step(self, nsteps=1):
        t = abi.FlyTallies()
        self._check(self.lib.emc_step(self.ctx, nsteps, C.byref(t)))
        return t

    def potential(self):
        p = np.zeros(self.mesh.nnode)
        self._check(self.lib.emc_potential_get(self.ctx, _ptr(p, C.c_double)))
        return p

    def timing(self):
        ms = np.zeros(5)
        it = C.c_int(0)
        self._check(self.lib.emc_step_timing(self.ctx, _ptr(ms, C.c_double), C.byref(it)))
        return ms, it.value

    def heat(self):
        nn = self.mesh.nnode
        he = np.zeros(nn, np.int64)
        hh = np.zeros(nn, np.int64)
        self._check(self.lib.emc_heat_get(self.ctx, _ptr(he, C.c_int64), _ptr(hh, C.c_int64)))
        return he, hh

    def currents(self):
        pny = self.mesh.pny
        r = np.zeros(pny, np.int64)
        lc = np.zeros(pny, np.int64)
        self._check(self.lib.emc_currents_get(self.ctx, _ptr(r, C.c_int64), _ptr(lc, C.c_int64)))
        return r, lc

    def tallies_reset(self):
        self._check(self.lib.emc_tallies_reset(self.ctx))

    # ---- statistics window and the data/ writers
    def new_window(self):
        """Host arrays of one statistics window (zeros) and the struct that points at them."""
        nn, nc, ny = self.mesh.nnode, self.mesh.ncell, self.mesh.pny
        w = {"stat_pot": np.zeros(nn), "stat_qc_pot": np.zeros(nn), "node_stat": np.zeros(10 * nn, np.int64),
             "cell_count": np.zeros(nc, np.int64), "heat_e": np.zeros(nn, np.int64),
             "heat_h": np.zeros(nn, np.int64), "rcur": np.zeros(ny, np.int64), "lcur": np.zeros(ny, np.int64)}
        st = abi.StatWindow()
        st.stat_pot = _ptr(w["stat_pot"], C.c_double)
        st.stat_qc_pot = _ptr(w["stat_qc_pot"], C.c_double)
        for k in ("node_stat", "cell_count", "heat_e", "heat_h", "rcur", "lcur"):
            setattr(st, k, _ptr(w[k], C.c_int64))
        w["struct"] = st
        return w

    def stat_window(self):
        """The device window: arrays, "steps" and "tallies" (abi.FlyTallies)."""
        w = self.new_window()
        self._check(self.lib.emc_stat_window_get(self.ctx, C.byref(w["struct"])))
        w["steps"] = w["struct"].steps
        w["tallies"] = w["struct"].tallies
        return w

    def write_stat(self, window, data_dir, step, heat_file=False):
        """Window files from a host window (new_window / stat_window); arrays set to
        None are left out (their files are not written)."""
        st = abi.StatWindow()
        st.steps = int(window.get("steps", 0))
        keep = []
        for k in ("stat_pot", "stat_qc_pot"):
            if window.get(k) is not None:
                a = np.ascontiguousarray(window[k], np.float64)
                keep.append(a)
                setattr(st, k, _ptr(a, C.c_double))
        for k in ("node_stat", "cell_count", "heat_e", "heat_h", "rcur", "lcur"):
            if window.get(k) is not None:
                a = np.ascontiguousarray(window[k], np.int64)
                keep.append(a)
                setattr(st, k, _ptr(a, C.c_int64))
        if window.get("tallies") is not None:
            st.tallies = window["tallies"]
        self._check(self.lib.emc_write_stat(self.ctx, C.byref(st), int(bool(heat_file)),
                                            os.fsencode(data_dir), int(step)))

    def output_stat(self, data_dir, step):
        self._check(self.lib.emc_output_stat(self.ctx, os.fsencode(data_dir), int(step)))

    def write_restart(self, parts, n, data_dir, step):
        self._check(self.lib.emc_write_restart(self.ctx, C.byref(parts.struct()), int(n),
                                               os.fsencode(data_dir), int(step)))

    def output_for_restart(self, data_dir, step):
        self._check(self.lib.emc_output_for_restart(self.ctx, os.fsencode(data_dir), int(step)))

    def output_pvolume(self, data_dir):
        self._check(self.lib.emc_output_pvolume(self.ctx, os.fsencode(data_dir)))

    def qc_potential(self, pot):
        """Effective potential of a node potential (quantum correction) on the GPU."""
        pot = np.ascontiguousarray(pot, np.float64)
        out = np.zeros(self.mesh.nnode)
        self._check(self.lib.emc_qc_potential(self.ctx, _ptr(pot, C.c_double), _ptr(out, C.c_double)))
        return out

    def heat_begin(self):
        self._check(self.lib.emc_heat_begin(self.ctx))

    def heat_step(self, nsteps=1):
        t = abi.FlyTallies()
        self._check(self.lib.emc_heat_step(self.ctx, nsteps, C.byref(t)))
        return t


def tallies_dict(t):
    return {
        "contact": np.array([[t.contact[i][j] for j in range(abi.CT_NCOL)] for i in range(abi.MAX_CONTACT)],
                            dtype=np.int64),
        "drops": np.array(list(t.drops), dtype=np.int64),
        "n_catch": t.n_catch, "n_gen": t.n_gen, "n_phonon": t.n_phonon,
        "n_impurity": t.n_impurity, "n_events": t.n_events,
    }


class Batch:
    """Engines that share the mesh, stepped together (emc_batch_*): every kernel
    of the step is one launch for all of them; each engine ends exactly where
    stepping it alone would."""

    def __init__(self, engines):
        self.lib = load_library()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.ctx for e in self.engines])
        st = C.c_int(0)
        self.b = self.lib.emc_batch_create(arr, len(self.engines), C.byref(st))
        if st.value != 0:
            msg = self.lib.emc_batch_last_error(self.b).decode() if self.b else "bad arguments"
            self.lib.emc_batch_destroy(self.b)
            self.b = None
            raise EmcError(st.value, msg)

    def step(self, nsteps=1):
        """nsteps steps of every engine; one FlyTallies per engine."""
        acc = (abi.FlyTallies * len(self.engines))()
        st = self.lib.emc_batch_step(self.b, nsteps, acc)
        if st != 0:
            raise EmcError(st, self.lib.emc_batch_last_error(self.b).decode())
        return list(acc)

    def timing(self):
        ms = np.zeros(5)
        it = C.c_int(0)
        self.lib.emc_batch_timing(self.b, _ptr(ms, C.c_double), C.byref(it))
        return ms, it.value

    def close(self):
        if self.b:
            self.lib.emc_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_dict(r):
    return {"state": r.state, "status": r.status, "converged": bool(r.converged), "windows": r.windows,
            "id_mean": r.id_mean, "id_half_width": r.id_half_width,
            "id_windows": [r.id_windows[q] for q in range(min(r.windows, abi.REQ_MAX_WINDOWS))],
            "steps": r.steps, "particles": r.particles, "nodes": r.nodes, "footprint_bytes": r.footprint_bytes,
            "t_submit": r.t_submit, "t_start": r.t_start, "t_done": r.t_done, "error": r.error.decode()}


class Server:
    """The serving loop (emc_server_*): submit decks, they run concurrently on the
    GPU until their drain current reaches the requested precision."""

    def __init__(self, device=0, hbm_budget_gb=32.0, max_active=64, table_dir=TABLE_DIR):
        self.lib = load_library()
        self._table_dir = table_dir.encode()
        st = C.c_int(0)
        self.s = self.lib.emc_server_create(self._table_dir, device, hbm_budget_gb * 2.0 ** 30, max_active,
                                            C.byref(st))
        if st.value != 0:
            self.s = None
            raise EmcError(st.value, "emc_server_create")

    def submit(self, input_txt, seed=1, window_steps=1000, skip_windows=1, min_windows=4, max_windows=64,
               rel_tol=0.02, abs_tol=0.0, drain_contact=1):
        o = abi.RequestOpts()
        self.lib.emc_default_request(C.byref(o))
        o.seed, o.window_steps, o.skip_windows = seed, window_steps, skip_windows
        o.min_windows, o.max_windows, o.rel_tol = min_windows, max_windows, rel_tol
        o.abs_tol, o.drain_contact = abs_tol, drain_contact
        rid = C.c_int64(0)
        st = self.lib.emc_server_submit(self.s, os.fsencode(input_txt), C.byref(o), C.byref(rid))
        if st != 0:
            raise EmcError(st, self.lib.emc_server_last_error(self.s).decode())
        return rid.value

    def poll(self, rid):
        r = abi.RequestResult()
        st = self.lib.emc_server_poll(self.s, rid, C.byref(r))
        if st != 0:
            raise EmcError(st, "no such request")
        return request_dict(r)

    def wait(self, rid, timeout=-1.0):
        r = abi.RequestResult()
        st = self.lib.emc_server_wait(self.s, rid, timeout, C.byref(r))
        if st not in (0, abi.EMC_ERR_STATE):
            raise EmcError(st, "no such request")
        return request_dict(r)

    def close(self):
        if self.s:
            self.lib.emc_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_deck(out_dir, **kw):
    """Write FinFET decks (see decks/make_finfet.py)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_finfet", os.path.join(DECK_DIR, "make_finfet.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.write_decks(out_dir, **kw)
    return os.path.join(out_dir, "input.txt")


def read_current(path):
    """Windows of a data/current file: a list of {"step", "current": {icont: A},
    "phonon", "impurity"} in file order (the fields current.py and plot_current.py read)."""
    out = []
    icont = None
    with open(path, "r", encoding="utf-8") as f:
        for line in f:
            if line.startswith("step"):
                out.append({"step": int(line.split("=")[1]), "current": {}, "phonon": None, "impurity": None})
            elif line.startswith("icont"):
                icont = int(line.split("=")[1])
            elif line.startswith("Current") and out:
                out[-1]["current"][icont] = float(line.split("=")[1])
            elif line.startswith("times of phonon scatter") and out:
                out[-1]["phonon"] = int(line.split(":")[1])
            elif line.startswith("times of impurity scatter") and out:
                out[-1]["impurity"] = int(line.split(":")[1])
    return out


# two-sided 95 % Student-t quantiles, 1 .. 30 degrees of freedom
_T975 = (12.706, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365, 2.306, 2.262, 2.228, 2.201, 2.179, 2.160, 2.145, 2.131,
         2.120, 2.110, 2.101, 2.093, 2.086, 2.080, 2.074, 2.069, 2.064, 2.060, 2.056, 2.052, 2.048, 2.045, 2.042)


def _t975(df):
    return _T975[df - 1] if df <= 30 else 1.959963984540054 + 2.4 / df


def current_estimate(path, contacts=(0, 1), skip=1, level=0.95):
    """Drain-current estimate from data/current with a confidence interval.

    The point estimate is the one of finfet/scripts/current.py: the mean over the
    windows after the first `skip`, and over `contacts`, of |Current|.  Each window
    (stat_step steps) is one batch mean; the interval is the Student-t interval
    of the mean of the per-window values.  Returns a dict with "mean",
    "half_width", "windows" and "lag1" (lag-1 autocorrelation of the window
    series: near 0 when windows are long enough for the batches to be independent).
    """
    win = read_current(path)[skip:]
    x = np.array([np.mean([abs(w["current"][c]) for c in contacts]) for w in win], np.float64)
    n = len(x)
    if n == 0:
        raise ValueError("no window after the skipped ones")
    mean = float(x.mean())
    if n < 2:
        return {"mean": mean, "half_width": float("inf"), "windows": n, "lag1": float("nan")}
    sd = float(x.std(ddof=1))
    try:
        from scipy.stats import t as student_t
        q = float(student_t.ppf(0.5 + level / 2, n - 1))
    except ImportError:
        if level != 0.95:
            raise ValueError("without scipy only the 95 % interval is available")
        q = _t975(n - 1)
    d = x - mean
    lag1 = float((d[1:] * d[:-1]).sum() / (d * d).sum()) if n > 2 and (d * d).sum() > 0 else float("nan")
    return {"mean": mean, "half_width": q * sd / np.sqrt(n), "windows": n, "lag1": lag1}
