// Table and mesh layouts consumed by the flight, shared by the HIP kernels and
// the host-side sampler.  Pointers refer to device memory in kernels and to
// host memory on the host; scalars are normalised as in the reference.
#pragma once
#include <stdint.h>

#include "devmath.h"
#include "emc_const.h"

namespace emc {

// One tetrahedron's data read on every flight event, contiguous: the four
// inward face planes (nx, ny, nz, d) of Band::HESSTET, the group velocity and
// base energy of Band::VELOTET, the variable-Gamma bound gamtet, the top-corner
// energy and the neighbour codes NeibrTet.  Padded to 256 B and aligned to the
// 128 B cache line, so a record is exactly two lines (208 B unpadded straddle
// two or three).
struct alignas(128) TetRec {
  double n[4][4];
  double vt[4];
  double gam;
  double e3;
  int neib[4];
  int ibt;
  int pad[15];
};
static_assert(sizeof(TetRec) == 256, "TetRec layout");

// What one try of get_state's rejection loop reads, in one 64-byte line: corner
// energies, the FASTSURF area factors and |v| of the tetrahedron.
struct alignas(64) SurfRec {
  double e[4];
  double faclow, facmed, fachigh, vgt;
};

// Corner coordinates and energies of a tetrahedron for getK, in one 128-byte line.
struct alignas(128) CornerRec {
  double x[4], y[4], z[4], e[4];
};

// A sphere enclosing a tetrahedron grown by the containment tolerance (every
// point that Flight::contains accepts lies inside it): the tetrahedron searches
// test it first and run the exact plane test only on its hits.
struct alignas(32) Sphere {
  double cx, cy, cz, r2;
};

// Point-location octree of one band's tetrahedra (built by build_tet_index).
struct TetIdxRoot {
  double lo[3], hi[3];  // closed box holding the band's grown tetrahedra
  int node, pad;        // root node
};
// count >= 0: leaf, entries first .. first+count-1; count < 0: inner node, children
// first .. first+7 (child bit 0: x >= mid, bit 1: y >= mid, bit 2: z >= mid)
struct TetIdxNode {
  int first, count;
};
// A tetrahedron of a leaf: the box of its grown tetrahedron widened to float (every
// point Flight::contains accepts lies in it) and its bucket ranges, packed: tlist bins
// lo = bits 0-8, hi = bits 9-17, tlists bins slo = bits 18-23, shi = bits 24-29 (an
// empty range has lo > hi).
struct alignas(32) TetIdxEntry {
  float lo[3], hi[3];
  int t;
  uint32_t bins;
};
static_assert(sizeof(TetIdxEntry) == 32, "TetIdxEntry layout");

struct BandDev {
  const TetRec* tr;
  const TetIdxRoot* troot;  // [MNB] or null: no index, Getisym scans the lists
  const TetIdxNode* tnode;
  const TetIdxEntry* tent;
  const Sphere* tsph;                         // [nt]
  const Sphere *tlist_s, *tlists_s, *slist_s; // aligned with tlist, tlists, slist
  const SurfRec* sr;
  const CornerRec* cr;
  const int* tet;  // [4][nt]
  const double *xkk, *ykk, *zkk, *eek;
  const double *vgt, *faclow, *facmed, *fachigh;
  const int *matsym, *indmat, *newisym;
  const double *xbz, *ybz, *zbz;
  const int *ntlist, *ptlist, *tlist, *nslist, *pslist, *slist, *ntlists, *ptlists, *tlists;
  const int* ibt;
  const double *energy, *dostetmax, *sumscatt, *scatte, *dose, *scattiie, *scatth, *dosh, *scattiih;
  const double *dos, *sumdos;  // [MNB][NTAB1], [NTAB1][3]: the initial-state sampler's DOS
  double dosmax[3];
  int nt, scpre, scprh, nbandE, nbandH, bandofE, bandofH, pad0;
  double dtable, emax, dlist, dlists, a0pi, sieg, mell, melt, meld;
  double difpr[3];
  double temptag, templag, templog, temptaf, templaf, temptof, temphop;
  const double* ptab[PTAB_N];  // dispersive phonon tables, column-contiguous; null: not used
};

struct MeshDev {
  const double *lx, *ly, *lz, *dx, *dy, *dz;
  const uint8_t* motion;    // [ncell][6]
  const int* attached;      // [ncell]
  const uint8_t* insurf;    // [ncell]
  const int* nearest_surf;  // [ncell]
  const double* da;         // [ncell]
  const int* surface_dir;   // [nsurf]
  int cnx, cny, cnz, pnx, pny, pnz, ncontact, flag_surface, flag_simp, pad0;
  double dt, eV0, conc0, scrt0, eps_si, frickel;
  double charge_sign[3];
};

struct FieldsDev {
  const double *ex, *ey, *ez, *hx, *hy, *hz, *rho, *e_rough, *e_sph, *h_rough, *h_sph;
};

// Counter slots of the flight tallies (int64).
enum {
  CNT_DROP0 = 0,  // 9 drop / error reasons
  CNT_CATCH = 9,
  CNT_GEN = 10,
  CNT_PHONON = 11,
  CNT_IMPURITY = 12,
  CNT_EVENTS = 13,
  CNT_N = 16
};

struct TallyDev {
  unsigned long long* contact;   // [16][5]
  unsigned long long* counters;  // [CNT_N]
  unsigned long long *heat_e, *heat_h;  // [nnode] or null
  unsigned long long *rcur, *lcur;      // [pny]
};

// Device particle arrays (SoA).  pk packs type | i | j | k; PK_DEAD marks a hole.
struct PartDev {
  double *x, *y, *z, *kx, *ky, *kz, *energy, *charge;
  uint32_t* pk;
  int32_t *itet, *isym;
  uint64_t* key;
};

// A GENERATE clone waiting to fly the rest of its step.
struct CloneRec {
  double x, y, z, kx, ky, kz, energy, charge, left_time;
  uint64_t key;
  uint32_t pk;
  int32_t itet, isym, pad;
};

EMC_HD inline uint32_t pk_make(int type, int i, int j, int k) {
  return ((uint32_t)type << 31) | ((uint32_t)i << 21) | ((uint32_t)j << 10) | (uint32_t)k;
}
EMC_HD inline int pk_type(uint32_t p) { return (int)(p >> 31); }
EMC_HD inline int pk_i(uint32_t p) { return (int)((p >> 21) & 0x3FF); }
EMC_HD inline int pk_j(uint32_t p) { return (int)((p >> 10) & 0x7FF); }
EMC_HD inline int pk_k(uint32_t p) { return (int)(p & 0x3FF); }

}  // namespace emc
