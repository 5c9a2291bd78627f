// Kernel argument blocks and launchers (kernels.hip).  Plain structs of device
// pointers and scalars, passed by value.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "tables.h"

namespace emc {

// Descriptors of the band tables, the mesh and the cell fields, resident in
// device memory: the flight kernels read them through one pointer, so the
// compiler loads each table pointer where it is used instead of keeping some
// eighty of them in scalar registers across the event loop.
struct FlyConst {
  BandDev B;
  MeshDev M;
  FieldsDev F;
};

struct FlyArgs {
  const FlyConst* C;           // device memory
  TallyDev T;
  PartDev P;
  long long n, cap;            // live range [0,n) and array capacity
  uint64_t step;
  int flag_heat, pad0;
  CloneRec* clones;            // clone buffer of this step
  unsigned long long* nclones; // device counter
  unsigned long long clone_cap;
  long long* dead;             // slots freed by caught particles
  unsigned long long* ndead;   // device counter
  unsigned long long dead_cap;
  uint32_t* slow;              // [cap] particles k_fly_fast hands to k_fly_slow
  unsigned long long* nslow;   // device counter, zero before the flight
  long long wave_slots;        // waves the GPU runs at once (CUs x 4 SIMDs)
};

// Deposit tiles: TILE_I x TILE_J x TILE_K cells (i, j, k) per workgroup; a tile's
// node block spans (TILE_I+1)(TILE_J+1)(TILE_K+1) nodes.  Tile index
// t = (tj*ntx + ti)*ntz + tk follows the node order (k fastest, then i, then j).
constexpr int TILE_I = 4, TILE_J = 8, TILE_K = 4;
constexpr int TILE_NODES = (TILE_I + 1) * (TILE_J + 1) * (TILE_K + 1);
constexpr int TILE_CELLS = TILE_I * TILE_J * TILE_K;
constexpr int DEP_NQ = 11;  // node charge + 5 statistics x 2 carrier types

// Counting sort of the particles by tile (drops empty slots).
// The population P was left by the previous sort in tile segments: slots
// toff[t] .. toff[t+1]-1 held the particles of tile t, and slots from toff[ntile]
// on hold particles added since (clones).  A particle still in the tile of its
// segment keeps its place among the segment's stayers; the others (movers: they
// changed tile, or were added) go after the stayers of their new tile.  With
// seg_valid = 0 (the population was written some other way) every particle is
// a mover.
struct SortArgs {
  PartDev P;                  // source
  PartDev Q;                  // destination, tile order
  long long n;
  int ntx, nty, ntz, ntile;
  unsigned* tcount;           // [ntile] movers into each tile; zero on entry, left zero by k_tile_scan
  unsigned* toff;             // [ntile + 1] segments of the previous sort on entry, of this one after k_tile_scan
  unsigned* tcursor;          // [ntile] zeroed by k_tile_scan
  unsigned* tstay;            // [ntile] stayers of each tile (k_tile_hist)
  unsigned* toff_old;         // [ntile + 1] the segments of the previous sort (k_tile_scan; all 0 without)
  int seg_valid, pad;
};


struct RhsArgs {
  size_t nnode;
  const unsigned long long* node_q;
  const double *pvol, *pdop, *dirichlet, *corr;
  const uint8_t* kind;
  double* rhs;
  double* density;  // may be null
};

// statistic()'s per-step window sums (src/mcmodel.cpp:1189-1196, 1234), taken by
// the deposit after its sums: before the solve, so pot is the previous step's solution
struct StatArgs {
  size_t nnode, ncell;
  const double* pot;                        // the potential the step starts from
  double* stat_pot;                         // += pot
  const unsigned long long* cell_n;         // this step's particle count per cell
  unsigned long long* stat_enum;            // += cell_n
  unsigned long long *heat_e, *heat_h;      // may be null: no heat tallied
  unsigned long long *stat_heat_e, *stat_heat_h;  // += heat
  const double* qc_pot;                     // may be null: no quantum correction
  double* stat_qc_pot;                      // += qc_pot
  int zero_heat;                            // heat phase: heat = 0 after it is added
  int pad;
};

// Quantum-correction effective potential at the nodes of the quantum region
// (effective_potential, src/mcmodel.cpp:2337-2376), from tables built on the host.
struct QcArgs {
  long long nq;                             // nodes of the quantum region
  int pnx, pny, pnz;
  int wlen[3];                              // window length per axis (table row stride)
  const int* qnodes;                        // [nq] node indices
  const uint16_t* flag;                     // [nnode] init_p_mat flags
  const int* lo[3];                         // per axis: window start of each target index
  const int* hi[3];                         // per axis: window end (inclusive)
  const double* w[3];                       // per axis: erf differences [target][src - lo][shape]
  double eb;                                // normalised barrier
  const double* pot;                        // Poisson potential
  double* qc_pot;                           // effective potential (quantum nodes written)
};

struct DepositArgs {
  BandDev B;
  MeshDev M;
  PartDev P;                  // tile order
  size_t nnode, ncell;
  int ntx, nty, ntz, ntile;
  const unsigned* toff;       // [ntile + 1]
  unsigned long long* blk;    // [ntile][TILE_NODES][DEP_NQ] per-tile node sums
  unsigned long long *node_q, *node_stat, *cell_q, *cell_n;  // any may be null
  int stat_accumulate;        // 1: node_stat += this step's sums, 0: node_stat = sums
  int pad0;
  // fused into the node gather and the tile kernel's cell pass when set:
  RhsArgs rhs;                // the Poisson right-hand side (rhs.rhs null: not computed)
  StatArgs st;                // the window sums (st.stat_pot null: not taken)
};

struct FieldArgs {
  size_t ncell;
  int cnx, cny, cnz, pnx, pny, pnz;
  const double *pot, *dx, *dy, *dz, *cvol, *da;
  const int* eeff_dir;
  const unsigned long long* cell_q;  // null: rho is an input
  double *ex, *ey, *ez, *rho;
  double* rates;  // [4][ncell]: e_rough, e_sph, h_rough, h_sph; null = skip
  int flag_rough, flag_sphonon;
  double eps_si, eps_ox, Tn, T0, field0, em0, ec0, hq0, conc0, scrt0, sirho;
  double surf_delta, surf_ail, surf_XIph, surf_theta, surf_gama, surf_Nbmod, surf_Npf, surf_Pft, surf_Pfn,
      surf_Kpha;
  double mass_eff[2];
};

// CG on the eliminated 7-point operator (level 0 of the multigrid hierarchy).
// R is the storage type of the vectors and of the operator: double, or float for
// the FP32 bulk of the mixed-precision solve (arithmetic is FP64 either way;
// only the words in memory are narrower).
template <class R>
struct PcgArgsT {
  using real = R;
  size_t n;
  int pnx, pny, pnz, nblocks;
  const R *diag, *cx, *cy, *cz, *b;
  const uint8_t* kind;
  R *x, *r, *p, *p2, *q, *z;  // p and p2: search direction of even and odd iterations
  double *part_a, *part_pq, *part_rr, *part_rz0, *part_rz1;  // block partial sums: part_a [3*nblocks], the others [nblocks]
  double* scal;  // [0] r.r [1] r.z [2] b.b [3] iterations done [4] r.r at the tolerance
  int* done;     // set once r.r reaches scal[4]: every later CG / V-cycle kernel returns at once
  double rtol;
  int pad0;
  // FP32 preconditioner of the FP64 CG (null otherwise): the residual is also
  // written in FP32 to r32, the cycle's input, and z is read from z32, its output
  float* r32;
  const float* z32;
};
using PcgArgs = PcgArgsT<double>;

constexpr int MG_MAXL = 10;

template <class R>
struct MgLevelDevT {
  int nx, ny, nz, pad;
  size_t n;
  // levels >= 1: upper half of the symmetric 27-point Galerkin stencil in FP32,
  // [14][n]: entry 0 is the diagonal, entry o-13 the coupling to offset o of
  // st27 for o = 14..26; the coupling to offset o < 13 is entry 13-o of the
  // neighbour.  The preconditioner only needs an SPD operator close to A_l.
  const float* c14;
  const R* wdinv;   // [n] omega/diag, 0 at fixed nodes
  R *x, *t, *r, *b; // work vectors; b of level 0 is the CG residual
};
using MgLevelDev = MgLevelDevT<double>;

// One multigrid cycle (V(2,2) on level 0, V(1,1) below) of the hierarchy built
// by mg_build (multigrid.h).
template <class R>
struct MgArgsT {
  using real = R;
  MgLevelDevT<R> L[MG_MAXL];
  int nl, nc;                     // levels, nodes of the coarsest level
  const R* cinv;                  // [nc*nc] coarsest inverse, row major
  const R *diag, *cx, *cy, *cz;   // level-0 operator
  int* done;
  // CG iterations: the first kernel of the cycle also books the residual of the
  // iteration (r.r from part_rr, npart block partials) into scal and sets done
  // once r.r <= scal[4]; null in the cycle that starts the solve
  const double* part_rr;
  double* scal;
  int npart, pad;
  // the cycle's result z = M r and the block partials of r.z (CG grid)
  R* z;
  double* part_rz;
};
using MgArgs = MgArgsT<double>;

// Iterative refinement of the mixed-precision solve: r = b - A x in FP64, its
// FP32 copy is the right-hand side of the FP32 bulk solve, whose correction d is
// added to x in FP64.
struct RefineArgs {
  PcgArgs a;        // FP64: b, x, r (r = b - A x), part_a, scal
  float* r32;       // FP32 copy of r
  const float* d32; // FP32 correction
};

// Initial population on the device: particle s belongs to spec group g with
// start[g] <= s < start[g+1] (type|cell in gpk[g], charge gq[g]).
struct SampleArgs {
  const FlyConst* C;
  PartDev P;
  long long n, ngroup;
  const unsigned long long* start;  // [ngroup + 1]
  const uint32_t* gpk;
  const double* gq;
  unsigned long long seed;
  double tmp_const;                 // 1 - exp(-emax)
};

// MultipleRefresh: the population P is resampled per (cell, carrier type) into Q.
struct MrArgs {
  PartDev P, Q;
  long long n;                 // slots of P
  long long ngroup;            // 2 * cells: group g = 2*cell + type
  int cnx, cnz;
  const int *desired_e, *desired_h;  // [ncell] target counts
  unsigned long long step;
  // work arrays: [n] sort keys, slots; [ngroup] per-group data
  unsigned *gk, *gk2, *idx, *idx2;
  unsigned long long *sk, *sk2;
  unsigned *gstart, *gend;
  unsigned long long *cnt, *off;
  double* tot;
  int* keep;
  unsigned long long* nlive;   // live slots of P (set by k_mr_bounds)
  void* temp;                  // hipCUB storage
  size_t temp_bytes;
};

// One word-copy job per device of a batch (k_copy_u64_b).
struct CopyArgs {
  const void* src;                // null: write zeros
  unsigned long long* dst;
  size_t n;
  int src_u32, pad;               // 1: src holds 32-bit words (widened)
};

// Device tables of the argument blocks of a batch of K devices that share the
// mesh (one entry per device, filled by the host before each launch).
struct BatchTables {
  int K, pad;
  const SortArgs* sort;
  const DepositArgs* dep;
  const PcgArgs* pcg;
  const MgArgs *mg_start, *mg_even, *mg_odd;
  // the FP32 cycle's tables when the contexts solve with it (null otherwise)
  const MgArgsT<float> *fmg_start, *fmg_even, *fmg_odd;
  const FieldArgs* field;
  const FlyArgs* fly;
};

}  // namespace emc

emc::MgArgs emc_mg_start_args(const emc::PcgArgs& a, const emc::MgArgs& m);
emc::MgArgs emc_mg_iter_args(const emc::PcgArgs& a, const emc::MgArgs& m, int it);
emc::MgArgsT<float> emc_mg_start_args(const emc::PcgArgsT<float>& a, const emc::MgArgsT<float>& m);
emc::MgArgsT<float> emc_mg_start_args(const emc::PcgArgs& a, const emc::MgArgsT<float>& m);
emc::MgArgsT<float> emc_mg_iter_args(const emc::PcgArgs& a, const emc::MgArgsT<float>& m, int it);
emc::MgArgsT<float> emc_mg_iter_args(const emc::PcgArgsT<float>& a, const emc::MgArgsT<float>& m, int it);
hipError_t emc_launch_pcg_start(const emc::PcgArgsT<float>& a, const emc::MgArgsT<float>& m, hipStream_t s);
hipError_t emc_launch_pcg_iter(const emc::PcgArgsT<float>& a, const emc::MgArgsT<float>& m, int it, hipStream_t s);
// mixed precision: r = b - A x (FP64) with its norms into a.scal[0] (r.r) and
// a.scal[2] (b.b), and r32 = r; then x += d32
hipError_t emc_launch_refine_resid(const emc::RefineArgs& r, hipStream_t s);
hipError_t emc_launch_refine_add(const emc::RefineArgs& r, hipStream_t s);
hipError_t emc_launch_zero_f32(float* p, size_t n, hipStream_t s);
hipError_t emc_launch_pcg_start_b(const emc::BatchTables& t, const emc::MgArgs& shape, int nblocks, hipStream_t s);
hipError_t emc_launch_pcg_iter_b(const emc::BatchTables& t, const emc::MgArgs& shape, int nblocks, int it,
                                 hipStream_t s);
hipError_t emc_launch_tile_sort_b(const emc::BatchTables& t, int ntile, long long nappend_max, hipStream_t s);
hipError_t emc_launch_deposit_b(const emc::BatchTables& t, int ntile, size_t nnode, hipStream_t s);
hipError_t emc_launch_field_b(const emc::BatchTables& t, size_t ncell, hipStream_t s);
hipError_t emc_launch_fly_b(const emc::BatchTables& t, long long nmax, long long wave_slots, hipStream_t s);
hipError_t emc_launch_copy_b(const emc::CopyArgs* t, int K, size_t nwords, hipStream_t s);

hipError_t emc_launch_sample(const emc::SampleArgs& a, hipStream_t s);
size_t emc_mr_temp_bytes(long long n, long long ngroup);
hipError_t emc_launch_mr_plan(const emc::MrArgs& a, hipStream_t s);
hipError_t emc_launch_mr_write(const emc::MrArgs& a, hipStream_t s);
hipError_t emc_launch_fly(const emc::FlyArgs& a, hipStream_t s);
hipError_t emc_launch_fly_clones(const emc::FlyArgs& a, unsigned long long b, unsigned long long e, hipStream_t s);
hipError_t emc_launch_merge(const emc::FlyArgs& a, unsigned long long begin, unsigned long long end,
                            const unsigned long long* ndead, unsigned long long* ticket, hipStream_t s);
hipError_t emc_launch_clones_dev(const emc::FlyArgs& a, hipStream_t s);
// nappend: an upper bound of the movers outside the segments (slots from the
// previous sort's total on, or every slot without segments), for the grid
hipError_t emc_launch_tile_sort(const emc::SortArgs& a, long long nappend, hipStream_t s);
hipError_t emc_launch_deposit(const emc::DepositArgs& a, hipStream_t s);
hipError_t emc_launch_rhs(const emc::RhsArgs& a, hipStream_t s);
hipError_t emc_launch_qc_potential(const emc::QcArgs& a, hipStream_t s);
hipError_t emc_launch_field(const emc::FieldArgs& a, hipStream_t s);
hipError_t emc_launch_zero(unsigned long long* p, size_t n, hipStream_t s);
int emc_pcg_blocks(size_t n);
hipError_t emc_launch_pcg_start(const emc::PcgArgs& a, const emc::MgArgs& m, hipStream_t s);
hipError_t emc_launch_pcg_iter(const emc::PcgArgs& a, const emc::MgArgs& m, int it, hipStream_t s);
// FP64 CG preconditioned by the FP32 cycle (a.r32 and a.z32 set)
hipError_t emc_launch_pcg_start(const emc::PcgArgs& a, const emc::MgArgsT<float>& m, hipStream_t s);
hipError_t emc_launch_pcg_iter(const emc::PcgArgs& a, const emc::MgArgsT<float>& m, int it, hipStream_t s);
