/*
 * emc.h - C ABI of the MI355X full-band ensemble Monte Carlo engine.
 *
 * One GPU, one context per device description.  Plain C: opaque handle,
 * plain pointers and sizes, int status codes.  Host views returned by the
 * *_view calls stay valid until emc_destroy.
 *
 * Reference seams this ABI replaces (chenglinye18-crypto/C-MC_electron):
 *   emc_create          MeshQuantities::MeshQuantities + initialize()   src/mcmodel.cpp:26-146
 *                       (getInputData :3824, read_grid_file :4595, read_device_file :4190,
 *                        Band::IELEC src/Band.cpp:4271, init_cell_data :3187,
 *                        init_point_data :3272, init_poisson_matrix :2898)
 *   emc_particles_init  init_particle_data / select_kstate        src/particle_init.cpp:112-219,
 *                                                                  src/mcmodel.cpp:3006-3057
 *   emc_deposit         statistic + particle_to_density + compute_cell_charge
 *                                                                  src/mcmodel.cpp:1167-1408, 348-417, 2664-2712
 *   emc_poisson_solve   PoissonSolver::solve_poisson              include/PoissonSolver.h:6-18,
 *                                                                  src/PoissonSolver.cpp:70-78
 *   emc_compute_field   compute_rhs + pot2field(_for_hole) + GetSurfRoughnessPhononScRate
 *                                                                  src/mcmodel.cpp:2858-2893, 2732-2855,
 *                                                                  src/GetSurfRoughnessPhononScRate.cpp:4-71
 *   emc_fly_step        MeshQuantities::update_particle            src/mcmodel.cpp:1460-1505
 *                       (particle_fly :1874-2161 and callees; local_migrate :1615-1638)
 *   emc_step            density() + update_particle per step       src/mcmodel.cpp:424-471, 499-566
 *   emc_run_file        main(input.txt)                            src/main.cpp:3-61
 */
#ifndef EMC_H
#define EMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (extend the reference's error enum, include/utils.h:120-128) */
#define EMC_OK 0
#define EMC_ERR_ARG (-1)      /* bad argument or size                       */
#define EMC_ERR_IO (-2)       /* deck or table file missing / malformed      */
#define EMC_ERR_DECK (-3)     /* deck semantics (alignment, ranges)          */
#define EMC_ERR_HIP (-4)      /* HIP runtime error                            */
#define EMC_ERR_CAPACITY (-5) /* particle or clone capacity exceeded          */
#define EMC_ERR_NOGPU (-6)    /* no HIP device / kernels not loadable         */
#define EMC_ERR_STATE (-7)    /* call out of order                             */
#define EMC_ERR_SOLVER (-8)   /* Poisson solver did not converge              */

/* ---- fixed-point scales of the int64 tallies (value = round(x * scale)) */
#define EMC_FX_Q 281474976710656.0      /* 2^48: per-step node/cell charge sums      */
#define EMC_FX_STAT 4294967296.0        /* 2^32: window statistics                   */
#define EMC_FX_HEAT 1099511627776.0     /* 2^40: heat per node                       */
#define EMC_FX_CONTACT 1099511627776.0  /* 2^40: contact charge / energy, currents   */

/* ---- drop / error reasons counted by the flight (reference prints, then drops or exits) */
#define EMC_DROP_TF_ZERO 0     /* free-flight time below MY_ZERO   src/mcmodel.cpp:2002-2006 */
#define EMC_DROP_NEG_ENERGY 1  /* energy < 0 after drift            src/mcmodel.cpp:2048-2056 */
#define EMC_DROP_NO_TET 2      /* Getisym found no tetrahedron      src/ParticleUtils.cpp:617-624 */
#define EMC_DROP_LOOPS 3       /* event-loop cap (65536) reached                              */
#define EMC_DROP_OFF_MESH 4    /* PASS/GENERATE across the mesh edge (slab edge on one rank)  */
#define EMC_DROP_BAD_RULE 5    /* motion rule not defined            src/CellHit.cpp:185-190  */
#define EMC_ERR_AREA 6         /* FASTSURF area < 0                  src/Band.cpp:692-700     */
#define EMC_ERR_CONTACT 7      /* attached contact index < 1         src/CellHit.cpp:444-447  */
#define EMC_ERR_CLONE_FULL 8   /* clone buffer full: clone lost                               */
#define EMC_NREASON 9

/* ---- contact tally columns */
#define EMC_CT_NGEN 0
#define EMC_CT_QGEN 1
#define EMC_CT_EGEN 2
#define EMC_CT_NCATCH 3
#define EMC_CT_QCATCH 4
#define EMC_CT_NCOL 5
#define EMC_MAX_CONTACT 16

/* Host view of the full-band tables (layouts documented per field; all
 * energies, wave vectors and rates normalised as in the reference). */
typedef struct emc_band_view {
  int nk, nt, nbt;
  int nband[3], bandof[3];
  int scpre, scprh;
  double dtable, emax, dlist, dlists, a0pi, sieg;
  double mell, melt, meld;
  double difpr[3];
  double temptag, templag, templog, temptaf, templaf, temptof, temphop;
  double Ni;
  const double *xkk, *ykk, *zkk, *eek;  /* [nk]                                  */
  const int *tet;                       /* [4][nt] corner k-point, energy ascending */
  const int *neib;                      /* [4][nt] NeibrTet                        */
  const int *ibt;                       /* [nt]                                    */
  const double *dnt;                    /* [4 comp][4 face][nt] face normals, offset */
  const double *vt;                     /* [nt][4] vx vy vz e0                      */
  const double *vgt, *gamtet, *faclow, *facmedium, *fachigh, *maxaovtet; /* [nt] */
  const double *xkct, *ykct, *zkct;     /* [nt]                                    */
  const int *matsym;                    /* [6][48]                                 */
  const int *indmat;                    /* [729]                                   */
  const int *newisym;                   /* [5][48]                                 */
  const double *xbz, *ybz, *zbz;        /* [5][48]                                 */
  const int *ntlist, *ptlist, *tlist;   /* [325][8], [325][8], [len]               */
  const int *nslist, *pslist, *slist;   /* [325][8], [325][8], [len]               */
  const int *ntlists, *ptlists, *tlists;/* [60][8],  [60][8],  [len]               */
  int tlist_len, slist_len, tlists_len;
  const double *energy;                 /* [7501]                                  */
  const double *dos;                    /* [8][7501]                               */
  const double *sumdos;                 /* [7501][3]                               */
  const double *dostetmax;              /* [8][7501]                               */
  const double *sumscatt;               /* [7501][8]                               */
  const double *scatte;                 /* [14][4][4] process, final, initial      */
  const double *dose;                   /* [14][4][7501]                           */
  const double *scattiie;               /* [4][4][7501] final, initial, itab       */
  const double *scatth;                 /* [4][3][3]                               */
  const double *dosh;                   /* [4][3][7501]                            */
  const double *scattiih;               /* [3][3][7501]                            */
  double DOSMAX[3];
  const double *tsphere;               /* [nt][4] centre x y z and radius^2 of the sphere
                                           enclosing each tetrahedron grown by MY_ZERO
                                           (prefilter of the tetrahedron searches)   */
  int tindex_nodes, tindex_entries;    /* point-location octree answering Getisym's list
                                           scans (0 nodes: not built, the scans run)  */
  const void *tindex_root;             /* [8 bands] {double lo[3], hi[3]; int32 node, pad} */
  const int *tindex_node;              /* [nodes][2] {first, count}; count < 0: children
                                           first..first+7 (bit0 x, bit1 y, bit2 z upper) */
  const void *tindex_entry;            /* [entries] {float lo[3], hi[3]; int32 tet;
                                           uint32 bins: tlist lo 0-8, hi 9-17, tlists lo
                                           18-23, hi 24-29}                           */
  const double *ptab[6];               /* dispersive phonon tables (init_deep), [row][col]:
                                           kloem 201x120, kloab 201x110, ktoem 51x160,
                                           ktoab 51x150, klaem 51x160, klaab 51x150; NULL
                                           unless the deck sets Flag_QuantumCorrection    */
} emc_band_view;

/* Host view of the mesh.  cell c = (j*cnx + i)*cnz + k ; node n = (j*pnx + i)*pnz + k. */
typedef struct emc_mesh_view {
  int cnx, cny, cnz, pnx, pny, pnz;
  int64_t ncell, nnode;
  int ncontact, nsurf;
  double dt;                        /* normalised time step                  */
  double spr0, time0, eV0, pot0, conc0, scrt0, field0, curr0, velo0; /* units */
  double eps_si, eps_ox, frickel, Tn, T0, em0, hq0, ec0, sirho;
  double surf_delta, surf_ail, surf_XIph, surf_theta, surf_gama, surf_Nbmod, surf_Npf;
  double surf_Pft, surf_Pfn, surf_Kpha;
  double mass_eff[3], charge_sign[3];
  int flag_surface, flag_rough, flag_sphonon, flag_simp, flag_heat;
  const double *lx, *ly, *lz, *dx, *dy, *dz;
  const int *material, *attached, *nearest_surf, *eeff_dir; /* [ncell]           */
  const uint8_t *motion;            /* [ncell][6]                             */
  const uint8_t *insurf;            /* [ncell]                                */
  const double *cvol, *da;          /* [ncell]                                */
  const int *surface_dir;           /* [nsurf]                                */
  const double *pvol, *pdop, *vadd; /* [nnode]                                */
  const uint8_t *node_kind;         /* [nnode] 0 unknown, 1 contact, 2 identity */
  const double *dirichlet;          /* [nnode]                                */
  const double *diag, *cx, *cy, *cz;/* [nnode] eliminated SPD 7-point operator */
  const double *rhs_corr;           /* [nnode]                                */
  const double *stencil;            /* [nnode][7] operator as assembled        */
  const int32_t *mr_desired_e, *mr_desired_h; /* [ncell] MultipleRefresh targets:
                                       the deck's counts, raised in silicon to the
                                       initial count and the default_par_number
                                       floor (src/particle_init.cpp:44-70)       */
  /* quantum correction (src/mcmodel.cpp:2337-2662): Flag_QuantumCorrection,
     quantum_start_step, the ep_parm values (lengths and Eb normalised) and the
     quantumRegion cells (surface_scatter_range cells included, as in the reference) */
  int flag_qc, quantum_start_step;
  double qc_ratio[3], qc_theta[3], qc_eb;
  const uint8_t *qcregion;          /* [ncell]                                */
} emc_mesh_view;

/* Structure-of-arrays particle buffers (caller-owned on the host). */
typedef struct emc_particles {
  double *x, *y, *z, *kx, *ky, *kz, *energy, *charge;
  int32_t *i, *j, *k, *itet, *isym, *ptype;   /* ptype 0 electron, 1 hole; i<0 = empty */
  uint64_t *key;                              /* counter-based RNG key / identity      */
} emc_particles;

/* Per-cell inputs of the flight: fields, cell charge density Rho and the
 * surface rates, all normalised. */
typedef struct emc_cell_fields {
  double *ex, *ey, *ez;       /* electron field  (c_field_*)     */
  double *hx, *hy, *hz;       /* hole field      (c_h_field_*)   */
  double *rho;                /* c_par_charge after the 0.5*DA clamp */
  double *e_rough, *e_sph, *h_rough, *h_sph; /* surface roughness / phonon rates */
} emc_cell_fields;

/* Flight tallies (exact int64, order-independent). */
typedef struct emc_fly_tallies {
  int64_t contact[EMC_MAX_CONTACT][EMC_CT_NCOL];
  int64_t drops[EMC_NREASON];
  int64_t n_catch;           /* catch_par_num: caught + dropped            */
  int64_t n_gen;             /* gen_par                                      */
  int64_t n_phonon, n_impurity; /* real scattering events (sttt counters)    */
  int64_t n_events;          /* event-loop iterations                        */
  int64_t n_clone_passes;
  int64_t n_full_loop;       /* particles that needed the full event loop (the
                                rest drifted to the end of the step with no event) */
  int64_t n_mr_gen;          /* mr_gen_num: particles added (+) or removed (-) by
                                MultipleRefresh (emc_step, Flag_MultipleRefresh) */
} emc_fly_tallies;

/* Options for emc_create. */
typedef struct emc_opts {
  const char *table_dir;     /* directory with bs.si.asc and zd.si.asc          */
  int device;                /* HIP device ordinal; -1 = host only (no GPU)     */
  uint64_t seed;             /* run seed for the particle RNG keys              */
  double capacity_factor;    /* particle capacity = factor * initial count      */
  double poisson_rtol;       /* relative residual target of the PCG            */
  int poisson_maxit;
  int poisson_fp32;          /* precision of emc_step's and emc_poisson_solve's
                                solve: 0 FP64 CG with the FP64 multigrid cycle;
                                1 emc_step solves with emc_poisson_solve_mixed;
                                2 (default) FP64 CG with the cycle in FP32: the
                                CG vectors, residual and stopping test stay FP64,
                                the preconditioner streams half the bytes         */
  double poisson_inner_rtol; /* tolerance of the FP32 bulk solve (1e-5)          */
} emc_opts;

typedef struct emc_ctx emc_ctx;

void emc_default_opts(emc_opts *o);
emc_ctx *emc_create(const char *input_txt, const emc_opts *o, int *status);
void emc_destroy(emc_ctx *ctx);
const char *emc_last_error(emc_ctx *ctx);

int emc_band_view_get(emc_ctx *ctx, emc_band_view *out);
int emc_mesh_view_get(emc_ctx *ctx, emc_mesh_view *out);

/* MultipleRefresh (src/MultipleRefresh.cpp:9-143): per cell and carrier type, the
 * particles with |q| >= 1e-30 are kept when their count is within a factor 2 of
 * the cell's target (emc_mesh_view.mr_desired_*) and sum q^2 < 2 (sum q)^2 / n;
 * otherwise the target number of copies is drawn with probability |q| / sum|q|,
 * each with charge sum q / target.  Draws and keys are defined (the reference
 * uses its global ran1): the draws of cell c and type t at `step` are the Philox
 * stream of key mix64(0x4D554C5449524546 ^ mix64(2c + t + 1)), a cell's particles
 * are taken in ascending key order, copy i of parent p gets key
 * mix64(p ^ mix64(step + 0xD1B54A32D192ED03 (i + 1))).  The population is left in
 * cell order (electrons, then holes, of each cell).  emc_step calls it after the
 * flight of every mr_step-th step when the deck sets Flag_MultipleRefresh. */
int emc_multiple_refresh(emc_ctx *ctx, uint64_t step, int64_t *n_before, int64_t *n_after);

/* Checkpoint and resume (the reference's output_for_restart / Flag_restart,
 * src/mcmodel.cpp:853-965, in a format of this library): save writes the device
 * state a later load continues from exactly -- the live particles (compacted),
 * the potential (the next solve's warm start), the step counter, the window
 * statistics, heat and y-plane current accumulators.  Stepping a loaded context
 * gives the same particles, tallies and potential, bit for bit, as stepping the
 * saved one.  The file is binary, native byte order: "EMCCKPT1", u32 version 1,
 * i32 cnx cny cnz pad, u64 seed, u64 next step, i64 n, then the particle arrays
 * (x y z kx ky kz energy charge f64, packed cell u32, itet isym i32, key u64) and
 * the node/plane arrays. */
int emc_checkpoint_save(emc_ctx *ctx, const char *path);
int emc_checkpoint_load(emc_ctx *ctx, const char *path);

/* Getisym (src/ParticleUtils.cpp:395-625) on the host tables, the code the flight
 * kernels run: for each of n points k[q][0..2] (wave vector, normalised) of energy
 * energy[q] in band iband[q], isym[q] is the symmetry operation mapping k into the
 * irreducible wedge and itet[q] the tetrahedron holding it.  itet[q] is read first as
 * the tetrahedron to test before the search (the particle's current one).  itet[q]
 * is -1 where no tetrahedron holds the point (the reference's Flag_Catch).  Needs no
 * GPU. */
int emc_band_locate(emc_ctx *ctx, int64_t n, const double *k, const double *energy, const int *iband, int *itet,
                    int *isym);

/* get_state with getK (src/ParticleUtils.cpp:48-284) on the host tables, the code
 * the flight kernels run: n final-state draws, the q-th at energy energy[q] in
 * band iband[q] with symmetry operation isym[q], its random numbers from the
 * stream of (key[q], step).  itet[q] is read as the particle's tetrahedron and
 * receives the chosen one, k[3q .. 3q+2] the wave vector (normalised; 0 on a
 * self-scattering), ndraws[q] the random numbers used and selfsc[q] 1 where the
 * draw self-scatters (no list entry, or the 10000-try limit).  serial < 0 runs
 * the loop of get_state as written; serial >= 0 takes the tries after the first
 * `serial` in chunks of 128 draws, as the slow flight kernel's wave-parallel form
 * does (the same results by construction; this is its check).  No counterpart in
 * the reference: a seam for tests, like emc_band_locate.  Needs no GPU. */
int emc_band_final_state(emc_ctx *ctx, int64_t n, const uint64_t *key, uint64_t step, const double *energy,
                         const int *iband, const int *isym, double *k, int *itet, int *ndraws, int *selfsc,
                         int serial);

/* Initial population (host sampling as init_particle_data), uploaded to the GPU
 * when one is attached.  n_out receives the particle count. */
int emc_particles_init(emc_ctx *ctx, int64_t *n_out);
int emc_particles_count(emc_ctx *ctx, int64_t *n_out);
int emc_particles_upload(emc_ctx *ctx, const emc_particles *host, int64_t n);
int emc_particles_download(emc_ctx *ctx, emc_particles *host, int64_t cap, int64_t *n_out);

/* Cell fields used by the next emc_fly_step (replaces the field solve). */
int emc_fields_set(emc_ctx *ctx, const emc_cell_fields *host);
int emc_fields_get(emc_ctx *ctx, emc_cell_fields *host);

/* One flight step (update_particle): every particle flies dt from the state on
 * the GPU; GENERATE clones fly the rest of the step; caught particles leave.
 * flag_heat enables phonon heat tallies.  Tallies of this call go to *out. */
int emc_fly_step(emc_ctx *ctx, uint64_t step, int flag_heat, emc_fly_tallies *out);

/* Deposit (statistic + particle_to_density + compute_cell_charge) of the current
 * population.  Outputs (any may be NULL), int64 fixed point:
 *   node_q[nnode]       sum of q*w (EMC_FX_Q)                 particle_to_density
 *   node_stat[10][nnode] vxE vyE vzE |q|E E|q|E, then the hole five (EMC_FX_STAT)
 *   cell_q[3][ncell]    sum q electrons, sum q holes, sum |q| (EMC_FX_Q)
 *   cell_n[ncell]       particle count per cell (stat_enum)
 * The deposit first sorts the population on the device by 4x8x4-cell tile and
 * drops empty slots, so the order of emc_particles_download changes (the sums do
 * not: they are integer).                                                          */
int emc_deposit(emc_ctx *ctx, int64_t *node_q, int64_t *node_stat, int64_t *cell_q, int64_t *cell_n);

/* Poisson solve on the eliminated operator: rhs, sol of length nnode.  rhs holds
 * the reference right-hand side (compute_rhs); sol is the initial guess on entry.
 * FP64 CG preconditioned by one geometric multigrid cycle (V(2,2) on the mesh,
 * V(1,1) on the coarser levels), to ||r|| <= rtol ||b||; the cycle runs in FP32
 * when emc_opts.poisson_fp32 is 2 (the default), else in FP64. */
int emc_poisson_solve(emc_ctx *ctx, const double *rhs, double *sol, double rtol, int maxit,
                      int *iters, double *relres);

/* emc_poisson_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * solve's work vectors (staged rhs and sol, the CG vectors and partial sums,
 * the multigrid work vectors) live in one static 80 MiB device array per GPU,
 * held until the process ends: meshes up to about 1.07 M nodes with the FP32
 * cycle (9.8 doubles per node), 980 k with the FP64 cycle (10.6); a larger mesh
 * returns EMC_ERR_CAPACITY.  Same arguments, precision, results and iteration
 * counts as emc_poisson_solve.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU race, so the caller serialises
 * them. */
int emc_poisson_solve_noalloc(emc_ctx *ctx, const double *rhs, double *sol, double rtol, int maxit,
                              int *iters, double *relres);

/* The same solve in mixed precision: the multigrid-preconditioned CG runs with
 * FP32 vectors and operator (the FP32 bulk) to inner_rtol on the residual
 * equation, and FP64 iterative refinement (r = b - A x, x += d) repeats it until
 * the FP64 residual meets ||r|| <= rtol ||b|| (at most maxit_outer times).  The
 * result meets the same tolerance as emc_poisson_solve.  outer: refinement
 * steps, inner: FP32 CG iterations in all.  No counterpart in the reference
 * (KLU in FP64, src/PoissonSolver.cpp:48-78). */
int emc_poisson_solve_mixed(emc_ctx *ctx, const double *rhs, double *sol, double rtol, int maxit_outer,
                            double inner_rtol, int *outer, int *inner, double *relres);

/* Level `level` of the multigrid hierarchy used by emc_poisson_solve (no reference
 * counterpart: the reference factorises with KLU).  dims[3] = nodes along i, j, k;
 * a27 (may be NULL; levels >= 1) receives the [27][n] Galerkin coefficients, entry
 * ((dj+1)*3 + (di+1))*3 + (dk+1) coupling node n to n + (di, dj, dk); omega the
 * Jacobi damping.  Level 0 is the mesh view's 7-point operator.  EMC_ERR_ARG past
 * the last level.  Works without a GPU. */
int emc_poisson_mg_level(emc_ctx *ctx, int level, int dims[3], double *a27, double *omega);

/* A batch of devices stepped together.  Replaces running one process per bias
 * point (run(), src/mcmodel.cpp:498-566, with the bias in the deck's contact
 * blocks, src/mcmodel.cpp:4190-4584): the contexts share one GPU and the mesh
 * dimensions (the same geometry; contact voltages, particle counts and seeds may
 * differ), and every kernel of the step is one launch for all of them.  Each
 * context keeps its own state and results: n steps of the batch leave every
 * context exactly where n emc_step calls on it alone would (bit for bit).  No
 * quantum correction; the heating phase stays per context.  The contexts must
 * outlive the batch and are not stepped alone while it steps them.  They share
 * the solver options (poisson_fp32 0 or 2, not the mixed solve's 1).
 * acc (may be NULL) receives one tally block per context.  emc_batch_timing
 * gives the phase times (ms per step, as emc_step_timing) of the last call and
 * the largest CG iteration count of its last step. */
typedef struct emc_batch emc_batch;
emc_batch *emc_batch_create(emc_ctx *const *ctx, int n, int *status);
int emc_batch_step(emc_batch *b, int nsteps, emc_fly_tallies *acc);
int emc_batch_timing(emc_batch *b, double *ms5, int *pcg_iters_max);
const char *emc_batch_last_error(emc_batch *b);
void emc_batch_destroy(emc_batch *b);

/* Serving: requests (a deck, a seed and a target precision of the drain current)
 * arrive at any time and run concurrently on one GPU.  No counterpart in the
 * reference, which runs one deck per process (src/main.cpp:3-61) and leaves the
 * current's precision to finfet/scripts/current.py afterwards.  A worker thread
 * admits queued requests in arrival order while their estimated device-memory
 * footprint fits hbm_budget_bytes (a request always runs when nothing else
 * does) and at most max_active run, keeps their state resident on the GPU, and
 * steps every running request: those with the same mesh dimensions as one batch
 * (emc_batch), different meshes at the same time on their own streams.  A request
 * runs windows of window_steps steps; each window gives one drain current, the
 * Current output_stat prints for contact drain_contact (src/mcmodel.cpp:756-842;
 * contact 1, the drain of the reference's FinFET decks, whose finfet/Idrain.png
 * plots it), or with drain_contact < 0 the mean over contacts 0 and 1 of
 * |Current| (finfet/scripts/current.py).  After skip_windows windows it is done
 * when the 95 % Student-t half-width of the mean over the later windows is within
 * max(rel_tol |mean|, abs_tol) (and at least min_windows of them), or after
 * max_windows windows.  The
 * result depends only on the deck, seed and window settings, never on the other
 * requests: the same alone and in the mix, bit for bit.  Times are seconds since
 * emc_server_create. */
#define EMC_REQ_QUEUED 0
#define EMC_REQ_RUNNING 1
#define EMC_REQ_DONE 2
#define EMC_REQ_FAILED 3
#define EMC_REQ_MAX_WINDOWS 64
typedef struct emc_request_opts {
  uint64_t seed;
  int window_steps, skip_windows, min_windows, max_windows;
  double rel_tol;                 /* 0.02: +-2 % at 95 % confidence */
  double abs_tol;                 /* A; 0: relative precision only    */
  int drain_contact;              /* 1; < 0: current.py's mean of contacts 0, 1 */
  int pad0;
} emc_request_opts;
typedef struct emc_request_result {
  int state, status, converged, windows;
  double id_mean, id_half_width;  /* A, as data/current prints it              */
  double id_windows[EMC_REQ_MAX_WINDOWS];
  int64_t steps, particles, nodes;
  double footprint_bytes;         /* measured device memory of the request       */
  double t_submit, t_start, t_done;
  char error[256];
} emc_request_result;
typedef struct emc_server emc_server;
void emc_default_request(emc_request_opts *r);
emc_server *emc_server_create(const char *table_dir, int device, double hbm_budget_bytes, int max_active,
                              int *status);
int emc_server_submit(emc_server *s, const char *input_txt, const emc_request_opts *r, int64_t *id);
int emc_server_poll(emc_server *s, int64_t id, emc_request_result *out);
/* waits for the request to finish (timeout_s < 0: no limit); EMC_ERR_STATE on timeout */
int emc_server_wait(emc_server *s, int64_t id, double timeout_s, emc_request_result *out);
const char *emc_server_last_error(emc_server *s);
void emc_server_destroy(emc_server *s);

/* quantumPotential (src/mcmodel.cpp:2378-2405): the effective potential of a node
 * potential on the GPU; qc_pot[n] = pot[n] outside the quantum region.  Needs a
 * deck with Flag_QuantumCorrection (EMC_ERR_STATE otherwise).  emc_step applies it
 * itself from step quantum_start_step on: electrons fly in the field of the
 * corrected potential, holes in that of the Poisson potential. */
int emc_qc_potential(emc_ctx *ctx, const double *pot, double *qc_pot);

/* From a node potential: cell fields (pot2field) and surface rates, written
 * into the context's cell fields; cell_rho must hold the clamped Rho. */
int emc_compute_field(emc_ctx *ctx, const double *pot, const double *cell_rho);

/* Flight tallies kept on the device across steps: node heat (EMC_FX_HEAT,
 * heat_to_point src/heat.cpp:169-196) and per-y-plane crossing charge
 * (EMC_FX_CONTACT, rcurrent/lcurrent include/mcmodel.h:236).  Reset zeroes
 * them and the window statistics of emc_step (set_stat_zero src/mcmodel.cpp:716). */
int emc_heat_get(emc_ctx *ctx, int64_t *heat_e, int64_t *heat_h);
int emc_currents_get(emc_ctx *ctx, int64_t *rcur, int64_t *lcur);
int emc_tallies_reset(emc_ctx *ctx);

/* ---- statistics window and the data/ writers --------------------------------
 * Each emc_step / emc_heat_step adds to the window what statistic() adds
 * (src/mcmodel.cpp:1167-1408): the node statistics of emc_deposit, the
 * potential the step starts from (before its solve), the particle count of
 * every cell, the heat tallied so far (heat phase), and the step's contact and
 * scatter tallies.  emc_tallies_reset starts a new window (set_stat_zero). */
typedef struct emc_stat_window {
  int64_t steps;             /* steps added since the last reset                 */
  double *stat_pot;          /* [nnode] sum of the potential (stat_pot)           */
  double *stat_qc_pot;       /* [nnode] sum of the electrons' potential after the
                                quantum correction (stat_qc_pot); equal to
                                stat_pot when the correction is off            */
  int64_t *node_stat;        /* [10][nnode] as emc_deposit (EMC_FX_STAT)          */
  int64_t *cell_count;       /* [ncell] sum of the cell counts (stat_enum)        */
  int64_t *heat_e, *heat_h;  /* [nnode] sum of the node heat (EMC_FX_HEAT)        */
  int64_t *rcur, *lcur;      /* [pny] crossing charge (EMC_FX_CONTACT)            */
  emc_fly_tallies tallies;   /* contact columns and scatter counts of the window  */
} emc_stat_window;

/* Copy the device window into the non-null arrays of *w (and w->steps, w->tallies). */
int emc_stat_window_get(emc_ctx *ctx, emc_stat_window *w);

/* output_stat + current_scatter_info (src/mcmodel.cpp:971-1161, 756-842) from a
 * host window: data_dir/current (appended, truncated at step 0),
 * section_current<step>, heat<step> if heat_file, electron_num<step>,
 * pot<step>, qc_pot<step>, Electron<step>, Hole<step>, in the reference's
 * formats; the averages are over the deck's stat_step.  Keeps the window's
 * mean potential and carrier densities for the next heat file and for
 * emc_heat_begin.  Host only: works without a GPU.  NULL arrays skip their files. */
int emc_write_stat(emc_ctx *ctx, const emc_stat_window *w, int heat_file, const char *data_dir, int step);

/* emc_stat_window_get, then emc_write_stat; the heat file is written in the
 * heat phase (emc_heat_begin), as the reference's flag_heat. */
int emc_output_stat(emc_ctx *ctx, const char *data_dir, int step);

/* output_for_restart (src/mcmodel.cpp:895-965): data_dir/par_info_restart<step>,
 * per cell "i j k |Qe| Ne |Qh| Nh" then "-1 -1 -1", from a host population
 * (emc_write_restart, host only) or from the device one (emc_output_for_restart). */
int emc_write_restart(emc_ctx *ctx, const emc_particles *host, int64_t n, const char *data_dir, int step);
int emc_output_for_restart(emc_ctx *ctx, const char *data_dir, int step);

/* print_p_data(p_volume, "pvolume", conc0) (src/mcmodel.cpp:513): data_dir/pvolume. */
int emc_output_pvolume(emc_ctx *ctx, const char *data_dir);

/* Heating phase (compute_heat, src/heat.cpp:108-167): the field comes from the
 * mean potential of the last window written by emc_write_stat and stays fixed;
 * each emc_heat_step deposits (window statistics, then the node heat is added
 * to the window and cleared), computes the surface rates, flies with phonon
 * heat tallies and runs MultipleRefresh every mr_step steps; no Poisson solve.
 * Heat steps count from 0 and draw from RNG counters disjoint from emc_step's. */
int emc_heat_begin(emc_ctx *ctx);
int emc_heat_step(emc_ctx *ctx, int nsteps, emc_fly_tallies *acc);

/* Poisson right-hand side of the eliminated system from a node charge sum
 * (particle_to_density normalisation + compute_rhs, src/mcmodel.cpp:407-416, 2858-2893). */
int emc_rhs_from_charge(emc_ctx *ctx, const int64_t *node_q, double *rhs_elim);

/* Full step(s) on the GPU: deposit -> rhs -> PCG -> field/rates -> fly. */
int emc_step(emc_ctx *ctx, int nsteps, emc_fly_tallies *acc);
int emc_potential_get(emc_ctx *ctx, double *pot);

/* Timers of the last emc_step call, milliseconds per step (HIP events on the
 * engine's stream): deposit+rhs, Poisson, field+rates, flight with clone passes
 * and merge, and the k_fly kernel alone; and PCG iterations of the last solve. */
int emc_step_timing(emc_ctx *ctx, double *ms5, int *pcg_iters);

/* Run the deck like `main input.txt` (MeshQuantities::run and compute_heat,
 * src/mcmodel.cpp:498-566, src/heat.cpp:108-167): data/pvolume, total_step
 * steps with the window files every stat_step steps, then heat_steps heat
 * steps when Flag_compute_heat, with the window files every stat_heat_step.
 * Writes into ./data, which must exist. */
int emc_run_file(const char *input_txt, const emc_opts *o);

#ifdef __cplusplus
}
#endif
#endif /* EMC_H */
