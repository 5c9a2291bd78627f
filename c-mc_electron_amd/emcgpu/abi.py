"""ctypes mirror of include/emc.h (structs, constants, prototypes)."""
import ctypes as C

EMC_OK = 0
EMC_ERR_ARG = -1
EMC_ERR_IO = -2
EMC_ERR_DECK = -3
EMC_ERR_HIP = -4
EMC_ERR_CAPACITY = -5
EMC_ERR_NOGPU = -6
EMC_ERR_STATE = -7
EMC_ERR_SOLVER = -8

FX_Q = 2.0 ** 48
FX_STAT = 2.0 ** 32
FX_HEAT = 2.0 ** 40
FX_CONTACT = 2.0 ** 40

DROP_NAMES = ["tf_zero", "neg_energy", "no_tet", "loops", "off_mesh", "bad_rule",
              "err_area", "err_contact", "err_clone_full"]
NREASON = 9
MAX_CONTACT = 16
CT_NCOL = 5

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_u8 = C.POINTER(C.c_uint8)
_i32 = C.POINTER(C.c_int32)
_u64 = C.POINTER(C.c_uint64)
_i64 = C.POINTER(C.c_int64)


class BandView(C.Structure):
    _fields_ = [
        ("nk", C.c_int), ("nt", C.c_int), ("nbt", C.c_int),
        ("nband", C.c_int * 3), ("bandof", C.c_int * 3),
        ("scpre", C.c_int), ("scprh", C.c_int),
        ("dtable", C.c_double), ("emax", C.c_double), ("dlist", C.c_double), ("dlists", C.c_double),
        ("a0pi", C.c_double), ("sieg", C.c_double),
        ("mell", C.c_double), ("melt", C.c_double), ("meld", C.c_double),
        ("difpr", C.c_double * 3),
        ("temptag", C.c_double), ("templag", C.c_double), ("templog", C.c_double),
        ("temptaf", C.c_double), ("templaf", C.c_double), ("temptof", C.c_double), ("temphop", C.c_double),
        ("Ni", C.c_double),
        ("xkk", _d), ("ykk", _d), ("zkk", _d), ("eek", _d),
        ("tet", _i), ("neib", _i), ("ibt", _i), ("dnt", _d), ("vt", _d),
        ("vgt", _d), ("gamtet", _d), ("faclow", _d), ("facmedium", _d), ("fachigh", _d), ("maxaovtet", _d),
        ("xkct", _d), ("ykct", _d), ("zkct", _d),
        ("matsym", _i), ("indmat", _i), ("newisym", _i),
        ("xbz", _d), ("ybz", _d), ("zbz", _d),
        ("ntlist", _i), ("ptlist", _i), ("tlist", _i),
        ("nslist", _i), ("pslist", _i), ("slist", _i),
        ("ntlists", _i), ("ptlists", _i), ("tlists", _i),
        ("tlist_len", C.c_int), ("slist_len", C.c_int), ("tlists_len", C.c_int),
        ("energy", _d), ("dos", _d), ("sumdos", _d), ("dostetmax", _d), ("sumscatt", _d),
        ("scatte", _d), ("dose", _d), ("scattiie", _d), ("scatth", _d), ("dosh", _d), ("scattiih", _d),
        ("DOSMAX", C.c_double * 3),
        ("tsphere", _d),
        ("tindex_nodes", C.c_int), ("tindex_entries", C.c_int),
        ("tindex_root", C.c_void_p), ("tindex_node", C.POINTER(C.c_int)), ("tindex_entry", C.c_void_p),
        ("ptab", _d * 6),
    ]


class MeshView(C.Structure):
    _fields_ = [
        ("cnx", C.c_int), ("cny", C.c_int), ("cnz", C.c_int), ("pnx", C.c_int), ("pny", C.c_int), ("pnz", C.c_int),
        ("ncell", C.c_int64), ("nnode", C.c_int64),
        ("ncontact", C.c_int), ("nsurf", C.c_int),
        ("dt", C.c_double),
        ("spr0", C.c_double), ("time0", C.c_double), ("eV0", C.c_double), ("pot0", C.c_double),
        ("conc0", C.c_double), ("scrt0", C.c_double), ("field0", C.c_double), ("curr0", C.c_double),
        ("velo0", C.c_double),
        ("eps_si", C.c_double), ("eps_ox", C.c_double), ("frickel", C.c_double), ("Tn", C.c_double),
        ("T0", C.c_double), ("em0", C.c_double), ("hq0", C.c_double), ("ec0", C.c_double), ("sirho", C.c_double),
        ("surf_delta", C.c_double), ("surf_ail", C.c_double), ("surf_XIph", C.c_double),
        ("surf_theta", C.c_double), ("surf_gama", C.c_double), ("surf_Nbmod", C.c_double),
        ("surf_Npf", C.c_double), ("surf_Pft", C.c_double), ("surf_Pfn", C.c_double), ("surf_Kpha", C.c_double),
        ("mass_eff", C.c_double * 3), ("charge_sign", C.c_double * 3),
        ("flag_surface", C.c_int), ("flag_rough", C.c_int), ("flag_sphonon", C.c_int),
        ("flag_simp", C.c_int), ("flag_heat", C.c_int),
        ("lx", _d), ("ly", _d), ("lz", _d), ("dx", _d), ("dy", _d), ("dz", _d),
        ("material", _i), ("attached", _i), ("nearest_surf", _i), ("eeff_dir", _i),
        ("motion", _u8), ("insurf", _u8),
        ("cvol", _d), ("da", _d),
        ("surface_dir", _i),
        ("pvol", _d), ("pdop", _d), ("vadd", _d),
        ("node_kind", _u8), ("dirichlet", _d),
        ("diag", _d), ("cx", _d), ("cy", _d), ("cz", _d),
        ("rhs_corr", _d), ("stencil", _d),
        ("mr_desired_e", C.POINTER(C.c_int32)), ("mr_desired_h", C.POINTER(C.c_int32)),
        ("flag_qc", C.c_int), ("quantum_start_step", C.c_int),
        ("qc_ratio", C.c_double * 3), ("qc_theta", C.c_double * 3), ("qc_eb", C.c_double),
        ("qcregion", _u8),
    ]


class Particles(C.Structure):
    _fields_ = [(n, _d) for n in ("x", "y", "z", "kx", "ky", "kz", "energy", "charge")] + \
               [(n, _i32) for n in ("i", "j", "k", "itet", "isym", "ptype")] + [("key", _u64)]


class CellFields(C.Structure):
    _fields_ = [(n, _d) for n in ("ex", "ey", "ez", "hx", "hy", "hz", "rho",
                                  "e_rough", "e_sph", "h_rough", "h_sph")]


class FlyTallies(C.Structure):
    _fields_ = [
        ("contact", (C.c_int64 * CT_NCOL) * MAX_CONTACT),
        ("drops", C.c_int64 * NREASON),
        ("n_catch", C.c_int64), ("n_gen", C.c_int64),
        ("n_phonon", C.c_int64), ("n_impurity", C.c_int64),
        ("n_events", C.c_int64), ("n_clone_passes", C.c_int64), ("n_full_loop", C.c_int64),
        ("n_mr_gen", C.c_int64),
    ]


class StatWindow(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("stat_pot", _d), ("stat_qc_pot", _d), ("node_stat", _i64), ("cell_count", _i64),
        ("heat_e", _i64), ("heat_h", _i64), ("rcur", _i64), ("lcur", _i64),
        ("tallies", FlyTallies),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("table_dir", C.c_char_p), ("device", C.c_int), ("seed", C.c_uint64),
        ("capacity_factor", C.c_double), ("poisson_rtol", C.c_double), ("poisson_maxit", C.c_int),
        ("poisson_fp32", C.c_int), ("poisson_inner_rtol", C.c_double),
    ]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
REQ_MAX_WINDOWS = 64


class RequestOpts(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("window_steps", C.c_int), ("skip_windows", C.c_int), ("min_windows", C.c_int),
        ("max_windows", C.c_int), ("rel_tol", C.c_double), ("abs_tol", C.c_double), ("drain_contact", C.c_int),
        ("pad0", C.c_int),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("state", C.c_int), ("status", C.c_int), ("converged", C.c_int), ("windows", C.c_int),
        ("id_mean", C.c_double), ("id_half_width", C.c_double), ("id_windows", C.c_double * REQ_MAX_WINDOWS),
        ("steps", C.c_int64), ("particles", C.c_int64), ("nodes", C.c_int64), ("footprint_bytes", C.c_double),
        ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double), ("error", C.c_char * 256),
    ]


PARTICLE_F64 = ("x", "y", "z", "kx", "ky", "kz", "energy", "charge")
PARTICLE_I32 = ("i", "j", "k", "itet", "isym", "ptype")
CELL_FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "rho", "e_rough", "e_sph", "h_rough", "h_sph")


def declare(lib):
    """Attach argtypes/restype to the exported functions of libemc.so."""
    P = C.POINTER
    ctx = C.c_void_p
    sig = {
        "emc_default_opts": (None, [P(Opts)]),
        "emc_create": (ctx, [C.c_char_p, P(Opts), P(C.c_int)]),
        "emc_destroy": (None, [ctx]),
        "emc_last_error": (C.c_char_p, [ctx]),
        "emc_band_view_get": (C.c_int, [ctx, P(BandView)]),
        "emc_mesh_view_get": (C.c_int, [ctx, P(MeshView)]),
        "emc_band_locate": (C.c_int, [ctx, C.c_int64, _d, _d, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "emc_band_final_state": (C.c_int, [ctx, C.c_int64, _u64, C.c_uint64, _d, P(C.c_int), P(C.c_int), _d,
                                           P(C.c_int), P(C.c_int), P(C.c_int), C.c_int]),
        "emc_multiple_refresh": (C.c_int, [ctx, C.c_uint64, P(C.c_int64), P(C.c_int64)]),
        "emc_checkpoint_save": (C.c_int, [ctx, C.c_char_p]),
        "emc_checkpoint_load": (C.c_int, [ctx, C.c_char_p]),
        "emc_particles_init": (C.c_int, [ctx, _i64]),
        "emc_particles_count": (C.c_int, [ctx, _i64]),
        "emc_particles_upload": (C.c_int, [ctx, P(Particles), C.c_int64]),
        "emc_particles_download": (C.c_int, [ctx, P(Particles), C.c_int64, _i64]),
        "emc_fields_set": (C.c_int, [ctx, P(CellFields)]),
        "emc_fields_get": (C.c_int, [ctx, P(CellFields)]),
        "emc_fly_step": (C.c_int, [ctx, C.c_uint64, C.c_int, P(FlyTallies)]),
        "emc_deposit": (C.c_int, [ctx, _i64, _i64, _i64, _i64]),
        "emc_poisson_solve": (C.c_int, [ctx, _d, _d, C.c_double, C.c_int, _i, _d]),
        "emc_poisson_solve_noalloc": (C.c_int, [ctx, _d, _d, C.c_double, C.c_int, _i, _d]),
        "emc_poisson_solve_mixed": (C.c_int, [ctx, _d, _d, C.c_double, C.c_int, C.c_double, _i, _i, _d]),
        "emc_poisson_mg_level": (C.c_int, [ctx, C.c_int, _i, _d, _d]),
        "emc_compute_field": (C.c_int, [ctx, _d, _d]),
        "emc_step": (C.c_int, [ctx, C.c_int, P(FlyTallies)]),
        "emc_potential_get": (C.c_int, [ctx, _d]),
        "emc_step_timing": (C.c_int, [ctx, _d, _i]),
        "emc_run_file": (C.c_int, [C.c_char_p, P(Opts)]),
        "emc_heat_get": (C.c_int, [ctx, _i64, _i64]),
        "emc_currents_get": (C.c_int, [ctx, _i64, _i64]),
        "emc_tallies_reset": (C.c_int, [ctx]),
        "emc_rhs_from_charge": (C.c_int, [ctx, _i64, _d]),
        "emc_stat_window_get": (C.c_int, [ctx, P(StatWindow)]),
        "emc_write_stat": (C.c_int, [ctx, P(StatWindow), C.c_int, C.c_char_p, C.c_int]),
        "emc_output_stat": (C.c_int, [ctx, C.c_char_p, C.c_int]),
        "emc_write_restart": (C.c_int, [ctx, P(Particles), C.c_int64, C.c_char_p, C.c_int]),
        "emc_output_for_restart": (C.c_int, [ctx, C.c_char_p, C.c_int]),
        "emc_output_pvolume": (C.c_int, [ctx, C.c_char_p]),
        "emc_heat_begin": (C.c_int, [ctx]),
        "emc_qc_potential": (C.c_int, [ctx, _d, _d]),
        "emc_heat_step": (C.c_int, [ctx, C.c_int, P(FlyTallies)]),
        "emc_batch_create": (C.c_void_p, [P(C.c_void_p), C.c_int, P(C.c_int)]),
        "emc_batch_step": (C.c_int, [C.c_void_p, C.c_int, P(FlyTallies)]),
        "emc_batch_timing": (C.c_int, [C.c_void_p, _d, _i]),
        "emc_batch_last_error": (C.c_char_p, [C.c_void_p]),
        "emc_batch_destroy": (None, [C.c_void_p]),
        "emc_default_request": (None, [P(RequestOpts)]),
        "emc_server_create": (C.c_void_p, [C.c_char_p, C.c_int, C.c_double, C.c_int, P(C.c_int)]),
        "emc_server_submit": (C.c_int, [C.c_void_p, C.c_char_p, P(RequestOpts), P(C.c_int64)]),
        "emc_server_poll": (C.c_int, [C.c_void_p, C.c_int64, P(RequestResult)]),
        "emc_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, P(RequestResult)]),
        "emc_server_last_error": (C.c_char_p, [C.c_void_p]),
        "emc_server_destroy": (None, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return sig
