// Kernel argument blocks and launchers (kernels.hip).  Plain structs of device
// pointers and scalars, passed by value.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "tables.h"

namespace emc {

struct FlyArgs {
  BandDev B;
  MeshDev M;
  FieldsDev F;
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
};

struct DepositArgs {
  BandDev B;
  MeshDev M;
  PartDev P;
  long long n;
  size_t nnode, ncell;
  unsigned long long *node_q, *node_stat, *cell_q, *cell_n;  // any may be null
};

struct RhsArgs {
  size_t nnode;
  const unsigned long long* node_q;
  const double *pvol, *pdop, *dirichlet, *corr;
  const uint8_t* kind;
  double* rhs;
  double* density;  // may be null
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

struct PcgArgs {
  size_t n;
  int pnx, pny, pnz, nblocks;
  const double *diag, *cx, *cy, *cz, *b;
  const uint8_t* kind;
  double *x, *r, *p, *q;
  double *part, *part2, *part3;  // [3*nblocks] each
  double* scal;                  // [0] r.r  [1] r.z  [2] b.b
};

}  // namespace emc

hipError_t emc_launch_fly(const emc::FlyArgs& a, hipStream_t s);
hipError_t emc_launch_fly_clones(const emc::FlyArgs& a, unsigned long long b, unsigned long long e, hipStream_t s);
hipError_t emc_launch_merge(const emc::FlyArgs& a, unsigned long long nclones, unsigned long long ndead,
                            unsigned long long* ticket, hipStream_t s);
hipError_t emc_launch_deposit(const emc::DepositArgs& a, hipStream_t s);
hipError_t emc_launch_rhs(const emc::RhsArgs& a, hipStream_t s);
hipError_t emc_launch_field(const emc::FieldArgs& a, hipStream_t s);
hipError_t emc_launch_zero(unsigned long long* p, size_t n, hipStream_t s);
int emc_pcg_blocks(size_t n);
hipError_t emc_launch_pcg_start(const emc::PcgArgs& a, hipStream_t s);
hipError_t emc_launch_pcg_iter(const emc::PcgArgs& a, int it, hipStream_t s);
