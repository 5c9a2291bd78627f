/*
 * oracle.h - CPU restatement of the reference's hot path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the product links, loads or calls this library.
 *
 * Each function cites the reference file:line it restates
 * (chenglinye18-crypto/C-MC_electron).  Pinning of each function (known-answer
 * vectors, closed forms, invariants) lives in tests/test_oracle_*.py.
 *
 * Defined semantics where the reference is undefined or serial-order dependent
 * (same definitions in the GPU path):
 *   - Random() is a counter-based Philox4x32-10 stream per particle and step,
 *     key = particle key, counter = (draw block, step lo, step hi, 0).
 *   - the surface-phonon branches draw drand = Random()*CALSCATTSUM(E, band)
 *     (src/ParSurfaceScatter.cpp:72,124 leave drand uninitialised).
 *   - SurfScGamma starts at 0 for every particle flight (a member in the reference).
 *   - sign(0) = +1 (src/utils.cpp:64 divides 0/0).
 *   - a GENERATE clone flies the rest of the step in which it is made and gets
 *     key mix64(parent ^ mix64(step + C*(ordinal+1))).
 *   - -log, sin and cos use the fdlibm methods restated in omath.c; pow(x,2) is
 *     x*x and pow(x,-2) is 1/(x*x).
 */
#ifndef EMC_ORACLE_H
#define EMC_ORACLE_H

#include <stdint.h>

#include "../include/emc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- math restatements (fdlibm methods) */
double orc_log(double x);
double orc_sin(double x);
double orc_cos(double x);

/* ---- RNG */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t orc_mix64(uint64_t z);
uint64_t orc_clone_key(uint64_t parent, uint64_t step, int ordinal);
uint64_t orc_init_key(uint64_t seed, uint64_t serial);
/* n uniform doubles of particle `key` at `step`, starting at draw 0 */
void orc_uniforms(uint64_t key, uint64_t step, int n, double *out);

/* ---- flight: update_particle / particle_fly (src/mcmodel.cpp:1460-2161).
 * Particles [0, n) fly one step; caught slots get i = -1.  Clones are
 * appended at n, n+1, ... (cap = capacity of the arrays); *n_out is the new
 * count.  heat_e/heat_h [nnode] (EMC_FX_HEAT), rcur/lcur [pny] (EMC_FX_CONTACT)
 * accumulate; they may be NULL when flag_heat is 0 / currents are unwanted. */
int orc_fly_step(const emc_band_view *b, const emc_mesh_view *m, const emc_cell_fields *f,
                 emc_particles *p, int64_t n, int64_t cap, uint64_t step, int flag_heat,
                 emc_fly_tallies *t, int64_t *heat_e, int64_t *heat_h, int64_t *rcur,
                 int64_t *lcur, int64_t *n_out);

/* ---- single-function probes used by the closed-form tests */
double orc_tet_time(const emc_band_view *b, int itet, int isym, const double k[3],
                    const double F[3], int *itetdir);
double orc_fastsurf(const emc_band_view *b, double e, int itet);
double orc_surf(const emc_band_view *b, double e, int itet);
/* get_state for (energy, band, isym) with particle key/step draws; returns
 * self-scatter flag, writes k, itet and the number of draws used. */
int orc_get_state(const emc_band_view *b, uint64_t key, uint64_t step, double energy, int iband,
                  int isym, double k[3], int *itet, int *ndraws);
int orc_getisym(const emc_band_view *b, const double k[3], double energy, int iband, int itet_hint,
                int *isym, int *itet);
/* EleImpScUpdateK (src/ElectronImpurityScatter.cpp:188-306) on one electron with
 * cell charge density rho; returns 1 and the new state when accepted, 0 when the
 * draw self-scatters. */
int orc_imp_update_k(const emc_band_view *b, const emc_mesh_view *m, uint64_t key, uint64_t step, double k[3],
                     int *isym, int *itet, double *energy, double rho);

/* ---- deposit: statistic + particle_to_density + compute_cell_charge
 * (src/mcmodel.cpp:1167-1408, 348-417, 2664-2712); outputs as emc_deposit. */
int orc_deposit(const emc_band_view *b, const emc_mesh_view *m, const emc_particles *p, int64_t n,
                int64_t *node_q, int64_t *node_stat, int64_t *cell_q, int64_t *cell_n);

/* ---- node charge density and Poisson right-hand side (src/mcmodel.cpp:407-416, 2858-2893)
 * rhs_ref: full system (contact rows = contact potential), rhs_elim: eliminated system. */
void orc_rhs(const emc_mesh_view *m, const int64_t *node_q, double *density, double *rhs_ref,
             double *rhs_elim);
/* ---- cell Rho clamp (src/mcmodel.cpp:2704-2711) */
void orc_cell_rho(const emc_mesh_view *m, const int64_t *cell_q, double *rho);
/* ---- pot2field (src/mcmodel.cpp:2776-2813) and surface rates
 * (src/GetSurfRoughnessPhononScRate.cpp:4-71) */
void orc_field(const emc_mesh_view *m, const double *pot, double *ex, double *ey, double *ez);
void orc_surface_rates(const emc_mesh_view *m, const double *ex, const double *ey, const double *ez,
                       const double *rho, double *e_rough, double *e_sph, double *h_rough,
                       double *h_sph);
/* ---- FP64 Jacobi-PCG on the eliminated operator; returns iterations */
int orc_pcg(const emc_mesh_view *m, const double *rhs_elim, double *x, double rtol, int maxit,
            double *relres);
/* ---- MultipleRefresh (src/MultipleRefresh.cpp:9-143) of a host population into `out`
 * (capacity cap); returns the new count or -1 when cap is too small.  Order and
 * draws as defined in omr.c (the same as the GPU path). */
int64_t orc_multiple_refresh(const emc_mesh_view *m, const emc_particles *in, int64_t n, const int32_t *desired_e,
                             const int32_t *desired_h, uint64_t step, emc_particles *out, int64_t cap);
uint64_t orc_mr_stream_key(int64_t cell, int type);
uint64_t orc_mr_child_key(uint64_t parent, uint64_t step, int64_t i);

/* y = A x with the operator as assembled by init_matrix (contact rows identity) */
void orc_apply_full(const emc_mesh_view *m, const double *x, double *y);

/* ---- quantum correction: out = pot, then the effective potential at the nodes
 * of the quantum region (src/mcmodel.cpp:2337-2662); returns 0 */
int orc_qc_potential(const emc_mesh_view *m, const double *pot, double *out);

/* ---- dispersive phonon final energy (test hook; FEKLOAB .. FEKTOEM,
 * src/ElectronPhononScatter.cpp:632-1022): process 5/6/9/10/11/12 at energy eet,
 * its one draw from the stream of (key, step) */
double orc_fek(const emc_band_view *b, int process, double eet, uint64_t key, uint64_t step);

#ifdef __cplusplus
}
#endif
#endif
