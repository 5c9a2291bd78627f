/*
 * ofly.c - CPU restatement of update_particle / particle_fly and every callee
 * on the flight path (test infrastructure only; see oracle.h for the defined
 * semantics).  One struct FS holds what MeshQuantities keeps in members.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

enum { NBE = 4, MWLE = 325, MWLES = 60, MTAB = 7500, NTAB1 = 7501, MNB = 8 };
enum { UP = 0, DOWN = 1, LEFT = 2, RIGHT = 3, FRONT = 4, BACK = 5 };
enum { PASS = 1, REFLECT = 2, SCATTOX = 3, PERIOD = 4, GENERATE = 5, CATCH = 6, CATCHGATE = 7, GENREF = 8 };
enum { PELEC = 0, PHOLE = 1 };
#define MY_ZERO 1e-13
#define PI 3.141592653589793238462643e0
#define LOOP_CAP 65536

#include "orng.h"

typedef struct {
  double x, y, z, kx, ky, kz, energy, charge, left_time;
  int i, j, k, itet, isym, ptype;
  uint64_t key;
} Clone;

typedef struct {
  const emc_band_view *b;
  const emc_mesh_view *m;
  const emc_cell_fields *f;
  /* particle */
  double x, y, z, kx, ky, kz, energy, charge, left_time;
  int icell, jcell, kcell, itet, isym, par_type, iband;
  uint64_t key;
  uint64_t step;
  /* flight */
  double vx, vy, vz, Ex, Ey, Ez, Rho, DA;
  double ImpScGamma, SurfScGamma, SurfRoughnessScRate, SurfPhononScRate;
  int itetdir, idir;
  int Flag_Catch, Flag_SelfScatter;
  int drop_reason;
  int nclone;
  int flag_heat;
  Rng rng;
  /* sinks */
  emc_fly_tallies *t;
  int64_t *heat_e, *heat_h, *rcur, *lcur;
  Clone *clones;
  int64_t nclones, clone_cap;
} FS;

#define B (s->b)
#define M (s->m)
#define TET(c, it) (B->tet[(size_t)(c) * B->nt + (it)])
#define DNT(comp, face, it) (B->dnt[((size_t)(comp) * 4 + (face)) * B->nt + (it)])
#define MATSYM(r, is) (B->matsym[(r) * 48 + (is)])
#define CELL(i, j, k) (((size_t)(j) * M->cnx + (i)) * M->cnz + (k))
#define NODE(i, j, k) (((size_t)(j) * M->pnx + (i)) * M->pnz + (k))

static double Random(FS *s) { return rng_next(&s->rng); }

static int64_t fx(double v, double scale) { return (int64_t)rint(v * scale); }

static int indmat_at(const emc_band_view *b, int sx, int sy, int sz, int p0, int p1, int p2) {
  return b->indmat[((((sx + 1) * 3 + (sy + 1)) * 3 + (sz + 1)) * 3 + p0) * 9 + p1 * 3 + p2];
}
static inline int sgn(double v) { return v < 0 ? -1 : 1; }

/* in_wedge / out_wedge, src/ParticleUtils.cpp:286-305, src/Band.cpp:555-570 */
static void in_wedge(FS *s, double *xo, double *yo, double *zo, double xi, double yi, double zi) {
  double v[3];
  v[MATSYM(0, s->isym) - 1] = xi * MATSYM(3, s->isym);
  v[MATSYM(1, s->isym) - 1] = yi * MATSYM(4, s->isym);
  v[MATSYM(2, s->isym) - 1] = zi * MATSYM(5, s->isym);
  *xo = v[0]; *yo = v[1]; *zo = v[2];
}
static void out_wedge_b(const emc_band_view *b, double *xo, double *yo, double *zo, double xi, double yi,
                        double zi, int isym) {
  const double v[3] = {xi, yi, zi};
  *xo = v[b->matsym[0 * 48 + isym] - 1] * b->matsym[3 * 48 + isym];
  *yo = v[b->matsym[1 * 48 + isym] - 1] * b->matsym[4 * 48 + isym];
  *zo = v[b->matsym[2 * 48 + isym] - 1] * b->matsym[5 * 48 + isym];
}
static void out_wedge(FS *s, double *xo, double *yo, double *zo, double xi, double yi, double zi) {
  out_wedge_b(B, xo, yo, zo, xi, yi, zi, s->isym);
}

/* GetV / GetEnergy, src/ParticleUtils.cpp:307-329 */
static void GetV(FS *s) {
  const double *v = &B->vt[4 * (size_t)s->itet];
  out_wedge(s, &s->vx, &s->vy, &s->vz, v[0], v[1], v[2]);
}
static void GetEnergy(FS *s) {
  double dx, dy, dz;
  in_wedge(s, &dx, &dy, &dz, s->kx, s->ky, s->kz);
  const int c0 = TET(0, s->itet);
  dx -= B->xkk[c0];
  dy -= B->ykk[c0];
  dz -= B->zkk[c0];
  const double *v = &B->vt[4 * (size_t)s->itet];
  s->energy = v[3] + v[0] * dx + v[1] * dy + v[2] * dz;
}

/* Band::FII / SURF / FASTSURF, src/Band.cpp:194-352, 665-703 */
static double FII(double x1, double y1, double z1, double x2, double y2, double z2, double x3, double y3,
                  double z3) {
  double sbx = (y2 - y1) * (z3 - z1) - (y3 - y1) * (z2 - z1);
  double sby = (x3 - x1) * (z2 - z1) - (x2 - x1) * (z3 - z1);
  double sbz = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
  return 0.50 * sqrt(sbx * sbx + sby * sby + sbz * sbz);
}

double orc_surf(const emc_band_view *b, double e, int it) {
  const int nt = b->nt;
  const int c0 = b->tet[it], c1 = b->tet[nt + it], c2 = b->tet[2 * (size_t)nt + it], c3 = b->tet[3 * (size_t)nt + it];
  const double *X = b->xkk, *Y = b->ykk, *Z = b->zkk;
  const double e0 = b->eek[c0], e1 = b->eek[c1], e2 = b->eek[c2], e3 = b->eek[c3];
  double r;
  if (!(e > e0 && e < e3)) return 0.0;
  if (e <= e1) {
    r = (e - e0) / (e1 - e0);
    double ax = r * (X[c1] - X[c0]), ay = r * (Y[c1] - Y[c0]), az = r * (Z[c1] - Z[c0]);
    r = (e - e0) / (e2 - e0);
    double bx = r * (X[c2] - X[c0]), by = r * (Y[c2] - Y[c0]), bz = r * (Z[c2] - Z[c0]);
    r = (e - e0) / (e3 - e0);
    double cx = r * (X[c3] - X[c0]), cy = r * (Y[c3] - Y[c0]), cz = r * (Z[c3] - Z[c0]);
    return FII(ax, ay, az, bx, by, bz, cx, cy, cz);
  } else if (e > e1 && e < e2) {
    r = (e - e0) / (e2 - e0);
    double ax = r * (X[c2] - X[c0]) + X[c0], ay = r * (Y[c2] - Y[c0]) + Y[c0], az = r * (Z[c2] - Z[c0]) + Z[c0];
    r = (e - e1) / (e2 - e1);
    double bx = r * (X[c2] - X[c1]) + X[c1], by = r * (Y[c2] - Y[c1]) + Y[c1], bz = r * (Z[c2] - Z[c1]) + Z[c1];
    r = (e - e0) / (e3 - e0);
    double cx = r * (X[c3] - X[c0]) + X[c0], cy = r * (Y[c3] - Y[c0]) + Y[c0], cz = r * (Z[c3] - Z[c0]) + Z[c0];
    r = (e - e1) / (e3 - e1);
    double dx = r * (X[c3] - X[c1]) + X[c1], dy = r * (Y[c3] - Y[c1]) + Y[c1], dz = r * (Z[c3] - Z[c1]) + Z[c1];
    double a1 = FII(ax, ay, az, bx, by, bz, cx, cy, cz);
    double a2 = FII(dx, dy, dz, bx, by, bz, cx, cy, cz);
    double a3 = FII(ax, ay, az, dx, dy, dz, cx, cy, cz);
    double a4 = FII(ax, ay, az, bx, by, bz, dx, dy, dz);
    return 0.50 * (a1 + a2 + a3 + a4);
  } else if (e > e2 && e < e3) {
    r = (e - e3) / (e0 - e3);
    double ax = r * (X[c0] - X[c3]), ay = r * (Y[c0] - Y[c3]), az = r * (Z[c0] - Z[c3]);
    r = (e - e3) / (e1 - e3);
    double bx = r * (X[c1] - X[c3]), by = r * (Y[c1] - Y[c3]), bz = r * (Z[c1] - Z[c3]);
    r = (e - e3) / (e2 - e3);
    double cx = r * (X[c2] - X[c3]), cy = r * (Y[c2] - Y[c3]), cz = r * (Z[c2] - Z[c3]);
    return FII(ax, ay, az, bx, by, bz, cx, cy, cz);
  }
  return 0.0;
}

static double fastsurf_core(const emc_band_view *b, double e, int it, int *bad) {
  const int nt = b->nt;
  const double e0 = b->eek[b->tet[it]], e1 = b->eek[b->tet[nt + it]], e2 = b->eek[b->tet[2 * (size_t)nt + it]],
               e3 = b->eek[b->tet[3 * (size_t)nt + it]];
  double r = 0.0;
  if (e > e0 && e < e3) {
    if (e <= e1)
      r = b->faclow[it] * (e - e0) * (e - e0);
    else if (e < e2) {
      if (e1 - e0 > 1e-10)
        r = b->faclow[it] * (e - e0) * (e - e0) - b->facmedium[it] * (e - e1) * (e - e1);
      else
        r = orc_surf(b, e, it);
    } else
      r = b->fachigh[it] * (e - e3) * (e - e3);
  }
  if (r < 0.0) { /* reference: "FASTSURF: Area<0" and exit(0) */
    *bad = 1;
    r = 0.0;
  }
  return r;
}
double orc_fastsurf(const emc_band_view *b, double e, int it) {
  int bad = 0;
  return fastsurf_core(b, e, it, &bad);
}
static double FASTSURF(FS *s, double e, int it) {
  int bad = 0;
  double r = fastsurf_core(B, e, it, &bad);
  if (bad) s->t->drops[EMC_ERR_AREA]++;
  return r;
}

/* Band interpolators, src/Band.cpp:1999-2031, 2098-2231, 2298-2321 */
static double CALSCATTSUM(const emc_band_view *b, double e, int ib) {
  if (e <= 0.0) return b->sumscatt[ib];
  if (e >= b->emax) return b->sumscatt[(size_t)MTAB * MNB + ib];
  int itab = (int)(e / b->dtable);
  double intp = (e - b->energy[itab]) / (b->energy[itab + 1] - b->energy[itab]);
  return b->sumscatt[(size_t)itab * MNB + ib] +
         intp * (b->sumscatt[(size_t)(itab + 1) * MNB + ib] - b->sumscatt[(size_t)itab * MNB + ib]);
}
static void CALSCATTE(const emc_band_view *b, double e, double *sca, int ib) {
  const int ini = ib - b->bandof[PELEC], sp = b->scpre;
#define SCE(is, f) b->scatte[((size_t)(is) * 4 + (f)) * 4 + ini]
#define DOSE(is, f, t) b->dose[((size_t)(is) * 4 + (f)) * NTAB1 + (t)]
#define IIE(f, t) b->scattiie[((size_t)(f) * 4 + ini) * NTAB1 + (t)]
  if (e <= 0e0) {
    for (int f = 0; f < b->nband[PELEC]; ++f) {
      for (int is = 0; is < sp - 1; ++is) sca[is + sp * f] = SCE(is, f) * DOSE(is, f, 0);
      sca[sp - 1 + sp * f] = IIE(f, 0);
    }
  } else if (e >= b->emax) {
    for (int f = 0; f < b->nband[PELEC]; ++f) {
      for (int is = 0; is < sp - 1; ++is) sca[is + sp * f] = SCE(is, f) * DOSE(is, f, MTAB);
      sca[sp - 1 + sp * f] = IIE(f, MTAB);
    }
  } else {
    int itab = (int)(e / b->dtable);
    double intp = (e - b->energy[itab]) / (b->energy[itab + 1] - b->energy[itab]);
    for (int f = 0; f < b->nband[PELEC]; ++f) {
      for (int is = 0; is < sp - 1; ++is)
        sca[is + sp * f] = SCE(is, f) * (DOSE(is, f, itab) + intp * (DOSE(is, f, itab + 1) - DOSE(is, f, itab)));
      sca[sp - 1 + sp * f] = IIE(f, itab) + intp * (IIE(f, itab + 1) - IIE(f, itab));
    }
  }
#undef SCE
#undef DOSE
#undef IIE
}
static void CALSCATTH(const emc_band_view *b, double e, double *sca, int ib) {
  const int ini = ib - b->bandof[PHOLE], sp = b->scprh;
#define SCH(is, f) b->scatth[((size_t)(is) * 3 + (f)) * 3 + ini]
#define DOSH(is, f, t) b->dosh[((size_t)(is) * 3 + (f)) * NTAB1 + (t)]
#define IIH(f, t) b->scattiih[((size_t)(f) * 3 + ini) * NTAB1 + (t)]
  if (e <= 0.0) {
    for (int f = 0; f < b->nband[PHOLE]; ++f) {
      for (int is = 0; is < sp - 1; ++is) sca[is + sp * f] = SCH(is, f) * DOSH(is, f, 0);
      sca[sp - 1 + sp * f] = IIH(f, 0);
    }
  } else if (e >= b->emax) {
    for (int f = 0; f < b->nband[PHOLE]; ++f) {
      for (int is = 0; is < sp - 1; ++is) sca[is + sp * f] = SCH(is, f) * DOSH(is, f, MTAB);
      sca[sp - 1 + sp * f] = IIH(f, MTAB);
    }
  } else {
    int itab = (int)(e / b->dtable);
    double intp = (e - b->energy[itab]) / (b->energy[itab + 1] - b->energy[itab]);
    for (int f = 0; f < b->nband[PHOLE]; ++f) {
      for (int is = 0; is < sp - 1; ++is)
        sca[is + sp * f] = SCH(is, f) * (DOSH(is, f, itab) + intp * (DOSH(is, f, itab + 1) - DOSH(is, f, itab)));
      sca[sp - 1 + sp * f] = IIH(f, itab) + intp * (IIH(f, itab + 1) - IIH(f, itab));
    }
  }
#undef SCH
#undef DOSH
#undef IIH
}
static double CALDOSTETMAX(const emc_band_view *b, double e, int ib) {
  const double *d = &b->dostetmax[(size_t)ib * NTAB1];
  if (e <= 0.0) return d[0];
  if (e >= b->emax) return d[MTAB];
  int itab = (int)(e / b->dtable);
  return d[itab] > d[itab + 1] ? d[itab] : d[itab + 1];
}
static void CONFBZ(const emc_band_view *b, double *x, double *y, double *z) {
  const double a = b->a0pi;
  for (int guard = 0; guard < 1024; ++guard) { /* same pass cap as the GPU path */
    int out = fabs(*x) > a || fabs(*y) > a || fabs(*z) > a || fabs(*x) + fabs(*y) + fabs(*z) > 1.50 * a;
    if (!out) return;
    *x = *x > 0.0 ? *x - a : *x + a;
    *y = *y > 0.0 ? *y - a : *y + a;
    *z = *z > 0.0 ? *z - a : *z + a;
  }
}

/* tetrahedron containment with tolerance, src/ParticleUtils.cpp:345-356 */
static int contains(const emc_band_view *b, int it, double xkl, double ykl, double zkl) {
  const size_t nt = b->nt;
#define D(c, f) b->dnt[((size_t)(c) * 4 + (f)) * nt + it]
  double d1 = D(0, 0) * xkl + D(1, 0) * ykl + D(2, 0) * zkl - D(3, 0);
  double d2 = D(0, 1) * xkl + D(1, 1) * ykl + D(2, 1) * zkl - D(3, 1);
  double d3 = D(0, 2) * xkl + D(1, 2) * ykl + D(2, 2) * zkl - D(3, 2);
  double d4 = D(0, 3) * xkl + D(1, 3) * ykl + D(2, 3) * zkl - D(3, 3);
#undef D
  return (d1 >= -MY_ZERO) && (d2 >= -MY_ZERO) && (d3 >= -MY_ZERO) && (d4 >= -MY_ZERO);
}

/* Getisym, src/ParticleUtils.cpp:395-625 (the cube list is empty, see band.cpp) */
static void Getisym(FS *s) {
  int pos[3] = {1, 1, 1}, mink, maxk;
  double tmin, tmax;
  if (fabs(s->kx) > fabs(s->ky)) { tmin = fabs(s->ky); tmax = fabs(s->kx); mink = 1; maxk = 0; }
  else { tmin = fabs(s->kx); tmax = fabs(s->ky); mink = 0; maxk = 1; }
  if (fabs(s->kz) < tmin) mink = 2;
  else if (fabs(s->kz) > tmax) maxk = 2;
  pos[mink] = 2;
  pos[maxk] = 0;
  s->isym = indmat_at(B, sgn(s->kx), sgn(s->ky), sgn(s->kz), pos[0], pos[1], pos[2]);
  double xkl, ykl, zkl;
  in_wedge(s, &xkl, &ykl, &zkl, s->kx, s->ky, s->kz);
  const int iband = s->iband;
  if (iband == B->ibt[s->itet] && contains(B, s->itet, xkl, ykl, zkl)) return;
  int ie = (int)(s->energy * B->dlists);
  if (ie < 0) ie = 0;
  if (ie < MWLES) {
    int p0 = B->ptlists[ie * MNB + iband], nn = B->ntlists[ie * MNB + iband];
    for (int itl = p0; itl <= p0 + nn - 1; ++itl) {
      int it = B->tlists[itl];
      if (iband == B->ibt[it] && contains(B, it, xkl, ykl, zkl)) { s->itet = it; return; }
    }
  }
  ie = (int)(s->energy * B->dlist);
  if (ie < 0) ie = 0;
  if (ie >= MWLE) ie = MWLE - 1;
  for (int pass = 0; pass < 3; ++pass) {
    int iq = pass == 0 ? ie : pass == 1 ? ie + 1 : ie - 1;
    if (pass == 1 && !(iq < MWLE)) continue;
    if (pass == 2 && !(iq >= 0)) continue;
    int p0 = B->ptlist[iq * MNB + iband], nn = B->ntlist[iq * MNB + iband];
    for (int itl = p0; itl <= p0 + nn - 1; ++itl) {
      int it = B->tlist[itl];
      if (iband == B->ibt[it] && contains(B, it, xkl, ykl, zkl)) { s->itet = it; return; }
    }
  }
  for (int it = 0; it < B->nt; ++it)
    if (iband == B->ibt[it] && contains(B, it, xkl, ykl, zkl)) { s->itet = it; return; }
  s->Flag_Catch = 1;
  s->drop_reason = EMC_DROP_NO_TET;
}

/* GetisymBZ, src/ParticleUtils.cpp:331-392 */
static void GetisymBZ(FS *s) {
  double xkl, ykl, zkl;
  in_wedge(s, &xkl, &ykl, &zkl, s->kx, s->ky, s->kz);
  const int iband = s->iband;
  if (iband == B->ibt[s->itet] && contains(B, s->itet, xkl, ykl, zkl)) return;
  int ie = (int)(s->energy * B->dlist);
  if (ie < 0) ie = 0;
  if (ie >= MWLE) ie = MWLE - 1;
  int p0 = B->pslist[ie * MNB + iband], nn = B->nslist[ie * MNB + iband];
  for (int itl = p0; itl <= p0 + nn - 1; ++itl) {
    int it = B->slist[itl];
    if (iband == B->ibt[it] && contains(B, it, xkl, ykl, zkl)) { s->itet = it; return; }
  }
  Getisym(s);
}

/* getK, src/ParticleUtils.cpp:122-284 */
static void getK(FS *s) {
  const int it = s->itet, nt = B->nt;
  const int c0 = TET(0, it), c1 = TET(1, it), c2 = TET(2, it), c3 = TET(3, it);
  const double *X = B->xkk, *Y = B->ykk, *Z = B->zkk, *E = B->eek;
  double rv1 = 1.0 - sqrt(Random(s));
  double rv2 = Random(s) * (1.0 - rv1);
  (void)nt;
  if ((s->energy < E[c0]) || (s->energy > E[c3])) s->energy = E[c0] + (E[c3] - E[c0]) * Random(s);
  const double e = s->energy;
  double r, xkl, ykl, zkl;
  if (e <= E[c1]) {
    r = (e - E[c0]) / (E[c1] - E[c0]);
    double ax = r * (X[c1] - X[c0]), ay = r * (Y[c1] - Y[c0]), az = r * (Z[c1] - Z[c0]);
    r = (e - E[c0]) / (E[c2] - E[c0]);
    double bx = r * (X[c2] - X[c0]), by = r * (Y[c2] - Y[c0]), bz = r * (Z[c2] - Z[c0]);
    r = (e - E[c0]) / (E[c3] - E[c0]);
    double cx = r * (X[c3] - X[c0]), cy = r * (Y[c3] - Y[c0]), cz = r * (Z[c3] - Z[c0]);
    xkl = rv1 * (bx - ax) + rv2 * (cx - ax) + ax + X[c0];
    ykl = rv1 * (by - ay) + rv2 * (cy - ay) + ay + Y[c0];
    zkl = rv1 * (bz - az) + rv2 * (cz - az) + az + Z[c0];
  } else if (e < E[c2]) {
    r = (e - E[c0]) / (E[c2] - E[c0]);
    double ax = r * (X[c2] - X[c0]) + X[c0], ay = r * (Y[c2] - Y[c0]) + Y[c0], az = r * (Z[c2] - Z[c0]) + Z[c0];
    r = (e - E[c1]) / (E[c2] - E[c1]);
    double bx = r * (X[c2] - X[c1]) + X[c1], by = r * (Y[c2] - Y[c1]) + Y[c1], bz = r * (Z[c2] - Z[c1]) + Z[c1];
    r = (e - E[c0]) / (E[c3] - E[c0]);
    double cx = r * (X[c3] - X[c0]) + X[c0], cy = r * (Y[c3] - Y[c0]) + Y[c0], cz = r * (Z[c3] - Z[c0]) + Z[c0];
    r = (e - E[c1]) / (E[c3] - E[c1]);
    double dx = r * (X[c3] - X[c1]) + X[c1], dy = r * (Y[c3] - Y[c1]) + Y[c1], dz = r * (Z[c3] - Z[c1]) + Z[c1];
    double a1 = FII(ax, ay, az, bx, by, bz, cx, cy, cz);
    double a2 = FII(dx, dy, dz, bx, by, bz, cx, cy, cz);
    double a3 = FII(ax, ay, az, dx, dy, dz, cx, cy, cz);
    double a4 = FII(ax, ay, az, bx, by, bz, dx, dy, dz);
    double ar = Random(s) * (a1 + a2 + a3 + a4);
    if (ar < a1) {
      xkl = rv1 * (bx - ax) + rv2 * (cx - ax) + ax;
      ykl = rv1 * (by - ay) + rv2 * (cy - ay) + ay;
      zkl = rv1 * (bz - az) + rv2 * (cz - az) + az;
    } else if (ar < (a1 + a2)) {
      xkl = rv1 * (bx - dx) + rv2 * (cx - dx) + dx;
      ykl = rv1 * (by - dy) + rv2 * (cy - dy) + dy;
      zkl = rv1 * (bz - dz) + rv2 * (cz - dz) + dz;
    } else if (ar < (a1 + a2 + a3)) {
      xkl = rv1 * (dx - ax) + rv2 * (cx - ax) + ax;
      ykl = rv1 * (dy - ay) + rv2 * (cy - ay) + ay;
      zkl = rv1 * (dz - az) + rv2 * (cz - az) + az;
    } else {
      xkl = rv1 * (bx - ax) + rv2 * (dx - ax) + ax;
      ykl = rv1 * (by - ay) + rv2 * (dy - ay) + ay;
      zkl = rv1 * (bz - az) + rv2 * (dz - az) + az;
    }
  } else {
    r = (e - E[c3]) / (E[c0] - E[c3]);
    double ax = r * (X[c0] - X[c3]), ay = r * (Y[c0] - Y[c3]), az = r * (Z[c0] - Z[c3]);
    r = (e - E[c3]) / (E[c1] - E[c3]);
    double bx = r * (X[c1] - X[c3]), by = r * (Y[c1] - Y[c3]), bz = r * (Z[c1] - Z[c3]);
    r = (e - E[c3]) / (E[c2] - E[c3]);
    double cx = r * (X[c2] - X[c3]), cy = r * (Y[c2] - Y[c3]), cz = r * (Z[c2] - Z[c3]);
    xkl = rv1 * (bx - ax) + rv2 * (cx - ax) + ax + X[c3];
    ykl = rv1 * (by - ay) + rv2 * (cy - ay) + ay + Y[c3];
    zkl = rv1 * (bz - az) + rv2 * (cz - az) + az + Z[c3];
  }
  out_wedge(s, &s->kx, &s->ky, &s->kz, xkl, ykl, zkl);
}

/* get_state, src/ParticleUtils.cpp:48-120 */
static void get_state(FS *s) {
  const double e = s->energy;
  const int iband = s->iband;
  if (e <= 0.0) {
    s->Flag_SelfScatter = 1;
    return;
  }
  double maxdostet = CALDOSTETMAX(B, e, iband);
  int smallfl, itab = (int)(e * B->dlists);
  if (itab < MWLES) smallfl = 1;
  else {
    itab = (int)(e * B->dlist);
    if (itab >= MWLE) itab = MWLE - 1;
    smallfl = 0;
  }
  int nted = smallfl ? B->ntlists[itab * MNB + iband] : B->ntlist[itab * MNB + iband];
  int base = smallfl ? B->ptlists[itab * MNB + iband] : B->ptlist[itab * MNB + iband];
  const int *list = smallfl ? B->tlists : B->tlist;
  if (nted == 0) {
    s->Flag_SelfScatter = 1;
    return;
  }
  int count2 = 0;
  for (;;) {
    s->itet = list[(int)(nted * Random(s)) + base];
    count2++;
    if ((count2 < 10000) && !((B->eek[TET(0, s->itet)] < e) && (B->eek[TET(3, s->itet)] > e))) continue;
    double dostet = FASTSURF(s, e, s->itet) / B->vgt[s->itet];
    if (((Random(s) * maxdostet) > dostet) && (count2 < 10000)) continue;
    break;
  }
  if (count2 >= 10000) s->Flag_SelfScatter = 1;
  else {
    s->Flag_SelfScatter = 0;
    getK(s);
  }
}

/* TALIND, src/ParticleUtils.cpp:626-748 */
static void TALIND(FS *s, int ityp) {
  int ital = 1;
  const double ax = fabs(s->kx), ay = fabs(s->ky), az = fabs(s->kz);
  if ((ax >= ay) && (ax >= az)) {
    if (s->kx >= 0.0) {
      if (ityp == 1) ital = 1;
      if (ityp == 2) ital = 2;
      if (ityp == 3) { const int t[5] = {0, 3, 4, 5, 6}; ital = t[(int)(Random(s) * 4.0 + 1.0)]; }
    }
    if (s->kx < 0.0) {
      if (ityp == 1) ital = 2;
      if (ityp == 2) ital = 1;
      if (ityp == 3) { const int t[5] = {0, 3, 4, 5, 6}; ital = t[(int)(Random(s) * 4.0 + 1.0)]; }
    }
  }
  if ((ay >= az) && (ay >= ax)) {
    if (s->ky >= 0.0) {
      if (ityp == 1) ital = 3;
      if (ityp == 2) ital = 4;
      if (ityp == 3) { const int t[5] = {0, 1, 2, 5, 6}; ital = t[(int)(Random(s) * 4.0 + 1.0)]; }
    }
    if (s->ky < 0.0) {
      if (ityp == 1) ital = 4;
      if (ityp == 2) ital = 3;
      if (ityp == 3) { const int t[5] = {0, 1, 2, 5, 6}; ital = t[(int)(Random(s) * 4.0 + 1.0)]; }
    }
  }
  if ((az > ay) && (az > ax)) {
    if (s->kz >= 0.0) {
      if (ityp == 1) ital = 5;
      if (ityp == 2) ital = 6;
      if (ityp == 3) { const int t[5] = {0, 1, 2, 3, 4}; ital = t[(int)(Random(s) * 4.0 + 1.0)]; }
    }
    if (s->kz < 0.0) {
      if (ityp == 1) ital = 6;
      if (ityp == 2) ital = 5;
      if (ityp == 3) { const int t[5] = {0, 1, 2, 3, 4}; ital = t[(int)(Random(s) * 4.0 + 1.0)]; }
    }
  }
  s->isym = (int)(Random(s) * 8.0) + 8 * (ital - 1);
}

/* heat_to_point, src/heat.cpp:169-196 */
static void heat_to_point(FS *s, double e, int64_t *heat) {
  if (!heat) return;
  const int i = s->icell, j = s->jcell, k = s->kcell;
  const double rx = (s->x - M->lx[i]) / M->dx[i];
  const double ry = (s->y - M->ly[j]) / M->dy[j];
  const double rz = (s->z - M->lz[k]) / M->dz[k];
  heat[NODE(i, j, k)] += fx(e * (1 - ry) * (1 - rx) * (1 - rz), EMC_FX_HEAT);
  heat[NODE(i + 1, j, k)] += fx(e * (1 - ry) * rx * (1 - rz), EMC_FX_HEAT);
  heat[NODE(i, j, k + 1)] += fx(e * (1 - ry) * (1 - rx) * rz, EMC_FX_HEAT);
  heat[NODE(i + 1, j, k + 1)] += fx(e * (1 - ry) * rx * rz, EMC_FX_HEAT);
  heat[NODE(i, j + 1, k)] += fx(e * ry * (1 - rx) * (1 - rz), EMC_FX_HEAT);
  heat[NODE(i + 1, j + 1, k)] += fx(e * ry * rx * (1 - rz), EMC_FX_HEAT);
  heat[NODE(i, j + 1, k + 1)] += fx(e * ry * (1 - rx) * rz, EMC_FX_HEAT);
  heat[NODE(i + 1, j + 1, k + 1)] += fx(e * ry * rx * rz, EMC_FX_HEAT);
}

/* The dispersive phonon final-energy draws used by EPSCAT in quantum-correction
 * runs, src/ElectronPhononScatter.cpp:632-1022, one function each as the
 * reference has them; tables indexed [row][col] as init_deep reads them
 * (src/ParticleUtils.cpp:750-808). */
#define PT(t, r, c, ncol) (B->ptab[t][(size_t)(r) * (ncol) + (c)])

static double FEKLOEM(FS *s, double eet) { /* :632-700, table kloem 201 x 120 */
  int i, j, k, it, ei;
  double ie, edeta, ran, ksum, efinal;
  ie = eet / B->templog;
  if (ie < 1.0) return 1.0e-10;
  if (ie < 3.05) { it = (int)((ie - 1.0) * 10.0 + 0.5); if (it < 1) it = 1; ei = it; edeta = 1.0; }
  else if (ie < 5.625) { it = (int)((ie - 3.0) * 4.0 + 0.5); if (it < 1) it = 1; ei = 20 + it; edeta = 2.0; }
  else if (ie < 13.25) { it = (int)((ie - 5.5) * 2.0 + 0.5); if (it < 1) it = 1; ei = 30 + it; edeta = 4.0; }
  else if (ie < 28.5) { it = (int)(ie - 13.0 + 0.5); if (it < 1) it = 1; ei = 45 + it; edeta = 8.0; }
  else { it = (int)((ie - 28.0) + 0.5); if (it < 1) it = 1; if (it > 60) it = 60; ei = 60 + it; edeta = 12.0; }
  ran = Random(s);
  i = 0;
  ksum = 0.0;
  while (ksum < ran) {
    i++;
    if (i == 201) { ksum = 1.1; i = 1; }
    ksum += PT(0, i - 1, ei - 1, 120);
  }
  j = i % 2;
  k = i / 2;
  if (j == 1) efinal = eet - B->templog * (1.0 - 0.01 * edeta * k);
  else efinal = eet - B->templog * (1.0 + 0.01 * edeta * k);
  return efinal;
}

static double FEKLOAB(FS *s, double eet) { /* :702-768, table kloab 201 x 110 */
  int i, j, k, it, ei;
  double ie, edeta, ran, ksum, efinal;
  ie = eet / B->templog;
  if (ie < 1.05) { it = (int)(ie * 10.0 + 0.5); if (it < 1) it = 1; ei = it; edeta = 1.0; }
  else if (ie < 3.625) { it = (int)((ie - 1.0) * 4.0 + 0.5); if (it < 1) it = 1; ei = 10 + it; edeta = 2.0; }
  else if (ie < 11.25) { it = (int)((ie - 3.5) * 2.0 + 0.5); if (it < 1) it = 1; ei = 20 + it; edeta = 4.0; }
  else if (ie < 26.5) { it = (int)((ie - 11.0) + 0.5); if (it < 1) it = 1; ei = 35 + it; edeta = 8.0; }
  else { it = (int)((ie - 26.0) + 0.5); if (it < 1) it = 1; if (it > 60) it = 60; ei = 50 + it; edeta = 12.0; }
  ran = Random(s);
  i = 0;
  ksum = 0.0;
  while (ksum < ran) {
    i++;
    ksum += PT(1, i - 1, ei - 1, 110);
    if (i == 201) { ksum = 1.1; i = 1; }
  }
  j = i % 2;
  k = i / 2;
  if (j == 1) efinal = eet + B->templog * (1.0 + 0.01 * edeta * k);
  else efinal = eet + B->templog * (1.0 - 0.01 * edeta * k);
  return efinal;
}

/* FEKLAAB (:770-826, klaab 51 x 150, phonon templaf) and FEKTOAB (:891-955, ktoab
 * 51 x 150, phonon temptof) share their bins; "it > 60" sets it to 115 in both. */
static double fek_ab51(FS *s, double eet, double hw, int tab, int clamp) {
  int i, j, k, it, ei;
  double ie, edeta, ran, ksum, efinal;
  ie = eet / hw;
  if (ie < 1.05) { it = (int)(ie * 10.0 + 0.5); if (it < 1) it = 1; ei = it; edeta = 0.25; }
  else if (ie < 3.625) { it = (int)((ie - 1.0) * 4.0 + 0.5); if (it < 1) it = 1; ei = 10 + it; edeta = 0.25; }
  else if (ie < 11.25) { it = (int)((ie - 3.5) * 2.0 + 0.5); if (it < 1) it = 1; ei = 20 + it; edeta = 0.5; }
  else { it = (int)((ie - 11.0) + 0.5); if (it < 1) it = 1; if (it > 60) it = 115; ei = 35 + it; edeta = 1.0; }
  ran = Random(s);
  i = 0;
  ksum = 0.0;
  while (ksum < ran) {
    i++;
    ksum += PT(tab, i - 1, ei - 1, 150);
    if (i == 51) { ksum = 1.1; i = 1; }
  }
  j = i % 2;
  k = i / 2;
  if (j == 1) efinal = eet + hw * (1.0 + 0.04 * edeta * k);
  else efinal = eet + hw * (1.0 - 0.04 * edeta * k);
  if (clamp && efinal <= 0.0) efinal = (1.0e-5) * hw;
  return efinal;
}

/* FEKLAEM (:828-889, klaem 51 x 160, templaf; below one phonon -1e-10) and FEKTOEM
 * (:957-1022, ktoem 51 x 160, temptof; below one phonon 1e-10). */
static double fek_em51(FS *s, double eet, double hw, int tab, int clamp, double below) {
  int i, j, k, it, ei;
  double ie, edeta, ran, ksum, efinal;
  ie = eet / hw;
  if (ie < 1.0) return below;
  if (ie < 3.05) { it = (int)((ie - 1) * 10.0 + 0.5); if (it < 1) it = 1; ei = it; edeta = 0.25; }
  else if (ie < 5.625) { it = (int)((ie - 3.0) * 4.0 + 0.5); if (it < 1) it = 1; ei = 20 + it; edeta = 0.5; }
  else if (ie < 13.25) { it = (int)((ie - 5.5) * 2.0 + 0.5); if (it < 1) it = 1; ei = 30 + it; edeta = 0.5; }
  else { it = (int)((ie - 13.0) + 0.5); if (it < 1) it = 1; if (it > 115) it = 115; ei = 45 + it; edeta = 1.0; }
  ran = Random(s);
  i = 0;
  ksum = 0.0;
  while (ksum < ran) {
    i++;
    ksum += PT(tab, i - 1, ei - 1, 160);
    if (i == 51) { ksum = 1.1; i = 1; }
  }
  j = i % 2;
  k = i / 2;
  if (j == 1) efinal = eet - hw * (1.0 - 0.04 * edeta * k);
  else efinal = eet - hw * (1.0 + 0.04 * edeta * k);
  if (clamp && efinal <= 0.0) efinal = (1.0e-5) * hw;
  return efinal;
}

static double FEKLAAB(FS *s, double eet) { return fek_ab51(s, eet, B->templaf, 5, 1); }
static double FEKLAEM(FS *s, double eet) { return fek_em51(s, eet, B->templaf, 4, 1, -1e-10); }
static double FEKTOAB(FS *s, double eet) { return fek_ab51(s, eet, B->temptof, 3, 0); }
static double FEKTOEM(FS *s, double eet) { return fek_em51(s, eet, B->temptof, 2, 0, 1e-10); }

/* EPSCAT, src/ElectronPhononScatter.cpp:98-195 (Jacoboni system; with the quantum
 * correction on, the optical and f-type LA/TO final energies come from the
 * dispersive tables, :148-163) */
static void EPSCAT(FS *s, int iscat) {
  const double eeold = s->energy, xkold = s->kx, ykold = s->ky, zkold = s->kz;
  const int ibold = s->iband, itetold = s->itet, isymold = s->isym;
  s->iband = iscat / B->scpre + B->bandof[s->par_type];
  if (M->flag_qc) switch (iscat % B->scpre) {
    case 0: TALIND(s, 1); break;
    case 1: s->energy = s->energy + B->temptag; TALIND(s, 2); break;
    case 2: s->energy = s->energy - B->temptag; TALIND(s, 2); break;
    case 3: s->energy = s->energy + B->templag; TALIND(s, 2); break;
    case 4: s->energy = s->energy - B->templag; TALIND(s, 2); break;
    case 5: s->energy = FEKLOAB(s, s->energy); TALIND(s, 2); break;
    case 6: s->energy = FEKLOEM(s, s->energy); TALIND(s, 2); break;
    case 7: s->energy = s->energy + B->temptaf; TALIND(s, 3); break;
    case 8: s->energy = s->energy - B->temptaf; TALIND(s, 3); break;
    case 9: s->energy = FEKLAAB(s, s->energy); TALIND(s, 3); break;
    case 10: s->energy = FEKLAEM(s, s->energy); TALIND(s, 3); break;
    case 11: s->energy = FEKTOAB(s, s->energy); TALIND(s, 3); break;
    case 12: s->energy = FEKTOEM(s, s->energy); TALIND(s, 3); break;
    default: break;
  }
  else switch (iscat % B->scpre) {
    case 0: TALIND(s, 1); break;
    case 1: s->energy = s->energy + B->temptag; TALIND(s, 2); break;
    case 2: s->energy = s->energy - B->temptag; TALIND(s, 2); break;
    case 3: s->energy = s->energy + B->templag; TALIND(s, 2); break;
    case 4: s->energy = s->energy - B->templag; TALIND(s, 2); break;
    case 5: s->energy = s->energy + B->templog; TALIND(s, 2); break;
    case 6: s->energy = s->energy - B->templog; TALIND(s, 2); break;
    case 7: s->energy = s->energy + B->temptaf; TALIND(s, 3); break;
    case 8: s->energy = s->energy - B->temptaf; TALIND(s, 3); break;
    case 9: s->energy = s->energy + B->templaf; TALIND(s, 3); break;
    case 10: s->energy = s->energy - B->templaf; TALIND(s, 3); break;
    case 11: s->energy = s->energy + B->temptof; TALIND(s, 3); break;
    case 12: s->energy = s->energy - B->temptof; TALIND(s, 3); break;
    default: break;
  }
  get_state(s);
  if (s->Flag_SelfScatter) {
    s->energy = eeold; s->kx = xkold; s->ky = ykold; s->kz = zkold;
    s->iband = ibold; s->itet = itetold; s->isym = isymold;
  }
  if (s->flag_heat) heat_to_point(s, (s->energy - eeold) * fabs(s->charge), s->heat_e);
}

/* ESCATII / HSCATII, src/ElectronPhononScatter.cpp:406-497, src/HolePhononScatter.cpp:164-234
 * (seciifl is false: no secondary particles) */
static void SCATII(FS *s, int iscat, int nproc) {
  const double eeold = s->energy, xkold = s->kx, ykold = s->ky, zkold = s->kz;
  const int ibold = s->iband, itetold = s->itet, isymold = s->isym;
  s->iband = iscat / nproc + B->bandof[s->par_type];
  s->energy = (s->energy - B->sieg) / 3.0;
  s->isym = ((int)(48 * Random(s)) % 48);
  get_state(s);
  if (s->Flag_SelfScatter) {
    s->energy = eeold; s->kx = xkold; s->ky = ykold; s->kz = zkold;
    s->iband = ibold; s->itet = itetold; s->isym = isymold;
  }
}

/* HPSCAT, src/HolePhononScatter.cpp:86-161 */
static void HPSCAT(FS *s, int iscat) {
  const double eeold = s->energy, xkold = s->kx, ykold = s->ky, zkold = s->kz;
  const int ibold = s->iband, itetold = s->itet, isymold = s->isym;
  s->iband = iscat / B->scprh + B->bandof[s->par_type];
  switch (iscat % B->scprh) {
    case 1: s->energy = s->energy + B->temphop; break;
    case 2: s->energy = s->energy - B->temphop; break;
    default: break;
  }
  s->isym = ((int)(48 * Random(s)) % 48);
  get_state(s);
  if (s->Flag_SelfScatter) {
    s->energy = eeold; s->kx = xkold; s->ky = ykold; s->kz = zkold;
    s->iband = ibold; s->itet = itetold; s->isym = isymold;
  }
  if (s->flag_heat) heat_to_point(s, (s->energy - eeold) * fabs(s->charge), s->heat_h);
}

/* process selection loop shared by the bulk and surface phonon routines */
static void select_process(FS *s, double drand) {
  double sca[14 * 4];
  const int elec = s->par_type == PELEC;
  const int np = elec ? B->scpre : B->scprh;
  const int nmax = np * B->nband[elec ? PELEC : PHOLE];
  if (elec) CALSCATTE(B, s->energy, sca, s->iband);
  else CALSCATTH(B, s->energy, sca, s->iband);
  int iscat = -1;
  double dscat = 0.0;
  while (drand > dscat) {
    iscat = iscat + 1;
    if (iscat >= nmax) {
      s->Flag_SelfScatter = 1;
      break;
    }
    dscat = dscat + sca[iscat];
  }
  if (s->Flag_SelfScatter) return;
  if (((iscat % np) + 1) == np) SCATII(s, iscat, np);
  else if (elec) EPSCAT(s, iscat);
  else HPSCAT(s, iscat);
}

/* ElectronPhononScatter / HolePhononScatter, src/ElectronPhononScatter.cpp:5-95,
 * src/HolePhononScatter.cpp:3-83 */
static void PhononScatter(FS *s) {
  const double scasum = CALSCATTSUM(B, s->energy, s->iband);
  const double g = B->gamtet[s->itet];
  double drand = scasum > g ? Random(s) * scasum : Random(s) * g;
  if (scasum < drand) {
    s->Flag_SelfScatter = 1;
    return;
  }
  s->Flag_SelfScatter = 0;
  select_process(s, drand);
}

/* GetImpScRate / GetImpScGamma, src/ElectronImpurityScatter.cpp:6-125 */
static double GetImpScRate(FS *s, double eel) {
  const double eps_si = M->eps_si, fr = M->frickel;
  if ((eel * M->eV0 < 0.120) && (s->DA * M->conc0 >= 1e22)) {
    double eebeta = fr * 4 * PI * s->Rho / (2.0 * B->meld * eps_si);
    size_t c = CELL(s->icell, s->jcell, s->kcell);
    if (M->flag_surface && M->flag_simp && M->insurf[c]) return 0.0; /* reference leaves it unset */
    double fr2 = (fr * s->Rho) * (fr * s->Rho);
    return s->DA * B->meld * sqrt(2.0 * B->meld * eel) / (PI * fr2 * (1.0 + 4.0 * eel * B->melt / (eebeta * B->meld)));
  }
  return 0.0;
}
static double GetImpScGamma(FS *s) {
  const int c0 = TET(0, s->itet), c3 = TET(3, s->itet);
  double eemax = M->frickel * 4 * PI * s->Rho / (8.0 * B->melt * M->eps_si);
  double gammax;
  if (eemax < B->eek[c0]) gammax = GetImpScRate(s, B->eek[c0]);
  else if (eemax > B->eek[c3]) gammax = GetImpScRate(s, B->eek[c3]);
  else gammax = GetImpScRate(s, eemax);
  double lo = 1.0 / M->scrt0;
  return gammax > lo ? gammax : lo;
}

/* EleImpScUpdateK, src/ElectronImpurityScatter.cpp:188-306 */
static void EleImpScUpdateK(FS *s) {
  double xkl, ykl, zkl;
  in_wedge(s, &xkl, &ykl, &zkl, s->kx, s->ky, s->kz);
  xkl = xkl - 0.85 * B->a0pi;
  xkl = xkl * sqrt(B->meld / B->mell);
  ykl = ykl * sqrt(B->meld / B->melt);
  zkl = zkl * sqrt(B->meld / B->melt);
  const double betaq = M->frickel * 4 * PI * s->Rho / M->eps_si;
  const double eebeta = betaq / (2.0 * B->meld);
  const double alfa = eebeta * B->meld / (2.0 * s->energy * B->melt);
  double pr = Random(s) / (1.0 + 0.5 * alfa);
  const double costr = 1.0 - alfa * pr / (1.0 - pr);
  const double sintr = sqrt(1.0 - costr * costr);
  pr = 2.0 * PI * Random(s);
  const double cospr = orc_cos(pr), sinpr = orc_sin(pr);
  const double ca = sqrt(xkl * xkl + ykl * ykl);
  const double cb = sqrt(ca * ca + zkl * zkl);
  const double xkll = cb * sintr * cospr, ykll = cb * sintr * sinpr, zkll = cb * costr;
  const double cosphi = xkl / ca, sinphi = ykl / ca, costheta = zkl / cb, sintheta = ca / cb;
  const double xklll = cosphi * costheta * xkll - sinphi * ykll + cosphi * sintheta * zkll;
  const double yklll = sinphi * costheta * xkll + cosphi * ykll + sinphi * sintheta * zkll;
  const double zklll = -sintheta * xkll + costheta * zkll;
  const double tn = betaq + B->mell / B->meld * (xkl - xklll) * (xkl - xklll) +
                    B->melt / B->meld * ((ykl - yklll) * (ykl - yklll) + (zkl - zklll) * (zkl - zklll));
  const double tm = betaq + B->melt / B->meld * (xkl - xklll) * (xkl - xklll) +
                    B->melt / B->meld * ((ykl - yklll) * (ykl - yklll) + (zkl - zklll) * (zkl - zklll));
  const double dn = 1.0 / (tn * tn), dnmax = 1.0 / (tm * tm);
  if (dn > Random(s) * dnmax) {
    xkl = xklll * sqrt(B->mell / B->meld);
    ykl = yklll * sqrt(B->melt / B->meld);
    zkl = zklll * sqrt(B->melt / B->meld);
    xkl = xkl + 0.85 * B->a0pi;
    out_wedge(s, &s->kx, &s->ky, &s->kz, xkl, ykl, zkl);
    CONFBZ(B, &s->kx, &s->ky, &s->kz);
    Getisym(s);
    GetEnergy(s);
    s->Flag_SelfScatter = 0;
  } else
    s->Flag_SelfScatter = 1;
}

/* ElectronImpurityScatter, src/ElectronImpurityScatter.cpp:128-185 (bhmrtfl false) */
static void ElectronImpurityScatter(FS *s) {
  double gamimp = GetImpScRate(s, s->energy);
  if (gamimp < Random(s) * s->ImpScGamma) s->Flag_SelfScatter = 1;
  else EleImpScUpdateK(s);
}

/* Reflect / Diffuse / Pass / Period, src/CellHit.cpp:214-432 */
static void mirror(FS *s, int axis) {
  int a = MATSYM(3, s->isym), b2 = MATSYM(4, s->isym), c = MATSYM(5, s->isym);
  if (axis == 0) { s->kx = -s->kx; a = -a; }
  else if (axis == 1) { s->ky = -s->ky; b2 = -b2; }
  else { s->kz = -s->kz; c = -c; }
  s->isym = indmat_at(B, a, b2, c, MATSYM(0, s->isym) - 1, MATSYM(1, s->isym) - 1, MATSYM(2, s->isym) - 1);
}
static void Reflect(FS *s) {
  if (s->idir == UP || s->idir == DOWN) mirror(s, 0);
  else if (s->idir == RIGHT || s->idir == LEFT) mirror(s, 1);
  else if (s->idir == FRONT || s->idir == BACK) mirror(s, 2);
}
static void Diffuse(FS *s) {
  const int ibandold = s->iband, itetold = s->itet, isymold = s->isym;
  const double xkold = s->kx, ykold = s->ky, zkold = s->kz;
  TALIND(s, 1);
  get_state(s);
  if (s->Flag_SelfScatter) {
    s->iband = ibandold; s->itet = itetold; s->isym = isymold;
    s->kx = xkold; s->ky = ykold; s->kz = zkold;
  }
  GetV(s);
  double v;
  int axis, flip;
  switch (s->idir) {
    case UP: v = s->vx; axis = 0; flip = v < 0.0; break;
    case RIGHT: v = s->vy; axis = 1; flip = v > 0.0; break;
    case DOWN: v = s->vx; axis = 0; flip = v > 0.0; break;
    case LEFT: v = s->vy; axis = 1; flip = v < 0.0; break;
    case FRONT: v = s->vz; axis = 2; flip = v < 0.0; break;
    case BACK: v = s->vz; axis = 2; flip = v > 0.0; break;
    default: return;
  }
  if (v == 0.0) {
    s->iband = ibandold; s->itet = itetold; s->isym = isymold;
    s->kx = xkold; s->ky = ykold; s->kz = zkold;
  }
  if (flip) mirror(s, axis);
}

static void contact_add(FS *s, int icont, int col, double v, int count) {
  if (icont < 0 || icont >= EMC_MAX_CONTACT || icont >= M->ncontact) {
    s->t->drops[EMC_ERR_CONTACT]++;
    return;
  }
  if (count) s->t->contact[icont][col] += 1;
  else s->t->contact[icont][col] += fx(v, EMC_FX_CONTACT);
}

static int neighbour(FS *s, int *ii, int *jj, int *kk) {
  *ii = s->icell; *jj = s->jcell; *kk = s->kcell;
  switch (s->idir) {
    case UP: (*ii)--; break;
    case DOWN: (*ii)++; break;
    case LEFT: (*jj)--; break;
    case RIGHT: (*jj)++; break;
    case FRONT: (*kk)--; break;
    case BACK: (*kk)++; break;
  }
  return *ii >= 0 && *ii < M->cnx && *jj >= 0 && *jj < M->cny && *kk >= 0 && *kk < M->cnz;
}

/* Generate, src/CellHit.cpp:435-501: clone on the other side of the plane */
static void Generate(FS *s) {
  s->t->n_gen++;
  int icont = M->attached[CELL(s->icell, s->jcell, s->kcell)] - 1;
  if (icont < 0) s->t->drops[EMC_ERR_CONTACT]++;
  else {
    contact_add(s, icont, EMC_CT_NGEN, 0, 1);
    contact_add(s, icont, EMC_CT_QGEN, s->charge, 0);
    contact_add(s, icont, EMC_CT_EGEN, s->energy, 0);
  }
  if (s->idir == RIGHT && s->rcur) s->rcur[s->jcell + 1] += fx(s->charge, EMC_FX_CONTACT);
  if (s->idir == LEFT && s->lcur) s->lcur[s->jcell] += fx(s->charge, EMC_FX_CONTACT);
  int ii, jj, kk;
  neighbour(s, &ii, &jj, &kk); /* in range: checked by HitCell */
  const int ord = s->nclone++;
  if (s->nclones >= s->clone_cap) {
    s->t->drops[EMC_ERR_CLONE_FULL]++;
    return;
  }
  Clone *c = &s->clones[s->nclones++];
  c->x = s->x; c->y = s->y; c->z = s->z;
  c->kx = s->kx; c->ky = s->ky; c->kz = s->kz;
  c->energy = s->energy; c->charge = s->charge; c->left_time = s->left_time;
  c->i = ii; c->j = jj; c->k = kk;
  c->itet = s->itet; c->isym = s->isym; c->ptype = s->par_type;
  c->key = orc_clone_key(s->key, s->step, ord);
}

/* CatchAtContact / CatchAtGate, src/CellHit.cpp:504-574 */
static void Catch(FS *s, int gate) {
  int ii, jj, kk;
  int icont = -1;
  if (neighbour(s, &ii, &jj, &kk)) icont = M->attached[CELL(ii, jj, kk)] - 1;
  if (icont < 0) {
    s->t->drops[EMC_ERR_CONTACT]++;
  }
  if (!gate) {
    if (s->idir == RIGHT && s->rcur) s->rcur[s->jcell + 1] += fx(s->charge, EMC_FX_CONTACT);
    if (s->idir == LEFT && s->lcur) s->lcur[s->jcell] += fx(s->charge, EMC_FX_CONTACT);
  }
  if (icont >= 0) {
    contact_add(s, icont, EMC_CT_NCATCH, 0, 1);
    contact_add(s, icont, EMC_CT_QCATCH, s->charge, 0);
    contact_add(s, icont, EMC_CT_EGEN, s->energy, 0); /* reference adds to EnergyGen */
  }
}

static void load_cell(FS *s) {
  size_t c = CELL(s->icell, s->jcell, s->kcell);
  s->Rho = s->f->rho[c];
  s->DA = M->da[c];
  const double sg = M->charge_sign[s->par_type];
  if (s->par_type == 0) {
    s->Ex = sg * s->f->ex[c]; s->Ey = sg * s->f->ey[c]; s->Ez = sg * s->f->ez[c];
  } else {
    s->Ex = sg * s->f->hx[c]; s->Ey = sg * s->f->hy[c]; s->Ez = sg * s->f->hz[c];
  }
}

/* HitCell, src/CellHit.cpp:96-211; returns 0 when the particle must leave the mesh */
static int HitCell(FS *s) {
  const int imot = M->motion[CELL(s->icell, s->jcell, s->kcell) * 6 + s->idir];
  int ii, jj, kk;
  const int inside = neighbour(s, &ii, &jj, &kk);
  if ((imot == PASS || imot == GENERATE || imot == GENREF) && !inside) return 0;
  if (imot == PASS) {
    switch (s->idir) {
      case UP: s->x = M->lx[s->icell]; s->icell--; break;
      case DOWN: s->x = M->lx[s->icell + 1]; s->icell++; break;
      case RIGHT:
        s->y = M->ly[s->jcell + 1];
        s->jcell++;
        if (s->rcur) s->rcur[s->jcell] += fx(s->charge, EMC_FX_CONTACT);
        break;
      case LEFT:
        s->y = M->ly[s->jcell];
        if (s->lcur) s->lcur[s->jcell] += fx(s->charge, EMC_FX_CONTACT);
        s->jcell--;
        break;
      case FRONT: s->z = M->lz[s->kcell]; s->kcell--; break;
      case BACK: s->z = M->lz[s->kcell + 1]; s->kcell++; break;
    }
  } else if (imot == REFLECT) {
    Reflect(s);
  } else if (imot == SCATTOX) {
    if (M->flag_surface) Reflect(s);
    else if (Random(s) > B->difpr[s->par_type]) Reflect(s);
    else Diffuse(s);
  } else if (imot == PERIOD) {
    if (s->idir == UP) s->x = M->lx[s->icell + 1];
    else if (s->idir == DOWN) s->x = M->lx[s->icell];
    else if (s->idir == LEFT) s->y = M->ly[s->jcell + 1];
    else if (s->idir == RIGHT) s->y = M->ly[s->jcell];
    else if (s->idir == FRONT) s->z = M->lz[s->kcell + 1];
    else if (s->idir == BACK) s->z = M->lz[s->kcell];
  } else if (imot == GENERATE || imot == GENREF) {
    Generate(s);
    Reflect(s);
  } else if (imot == CATCH) {
    Catch(s, 0);
    s->Flag_Catch = 1;
    s->drop_reason = -1;
  } else if (imot == CATCHGATE) {
    Catch(s, 1);
    s->Flag_Catch = 1;
    s->drop_reason = -1;
  } else {
    s->Flag_Catch = 1;
    s->drop_reason = EMC_DROP_BAD_RULE;
    return 1;
  }
  load_cell(s);
  return 1;
}

/* TetTime, src/TetHit.cpp:4-106 */
static double tet_time_core(const emc_band_view *b, int itet, int isym, double kx, double ky, double kz,
                            double Ex, double Ey, double Ez, int *itetdir) {
  double xkl, ykl, zkl, xfl, yfl, zfl, v[3];
  const int *ms = b->matsym;
  v[ms[0 * 48 + isym] - 1] = kx * ms[3 * 48 + isym];
  v[ms[1 * 48 + isym] - 1] = ky * ms[4 * 48 + isym];
  v[ms[2 * 48 + isym] - 1] = kz * ms[5 * 48 + isym];
  xkl = v[0]; ykl = v[1]; zkl = v[2];
  v[ms[0 * 48 + isym] - 1] = Ex * ms[3 * 48 + isym];
  v[ms[1 * 48 + isym] - 1] = Ey * ms[4 * 48 + isym];
  v[ms[2 * 48 + isym] - 1] = Ez * ms[5 * 48 + isym];
  xfl = v[0]; yfl = v[1]; zfl = v[2];
  const size_t nt = b->nt;
  double best = 0.0;
  for (int f = 0; f < 4; ++f) {
    const double n0 = b->dnt[((size_t)0 * 4 + f) * nt + itet], n1 = b->dnt[((size_t)1 * 4 + f) * nt + itet],
                 n2 = b->dnt[((size_t)2 * 4 + f) * nt + itet], n3 = b->dnt[((size_t)3 * 4 + f) * nt + itet];
    const double nf = n0 * xfl + n1 * yfl + n2 * zfl;
    double tfp;
    if (nf < 0) {
      tfp = (n3 - (n0 * xkl + n1 * ykl + n2 * zkl)) / nf;
      if (tfp < 0) tfp = 0;
    } else
      tfp = 1e99;
    if (f == 0 || tfp < best) {
      best = tfp;
      *itetdir = f;
    }
  }
  return best;
}
double orc_tet_time(const emc_band_view *b, int itet, int isym, const double k[3], const double F[3],
                    int *itetdir) {
  return tet_time_core(b, itet, isym, k[0], k[1], k[2], F[0], F[1], F[2], itetdir);
}

/* HitTet, src/TetHit.cpp:109-142 */
static void HitTet(FS *s) {
  const int nb = B->neib[(size_t)s->itetdir * B->nt + s->itet];
  if (nb >= 0) s->itet = nb;
  else if (nb == -5 || nb == -6) {
    const int ii = -nb - 2;
    s->kx = s->kx + B->xbz[ii * 48 + s->isym];
    s->ky = s->ky + B->ybz[ii * 48 + s->isym];
    s->kz = s->kz + B->zbz[ii * 48 + s->isym];
    s->isym = B->newisym[ii * 48 + s->isym];
    GetisymBZ(s);
    if (s->Flag_Catch) return;
    GetEnergy(s);
  } else
    s->isym = B->newisym[(-nb - 2) * 48 + s->isym];
}

/* CellTime, src/CellHit.cpp:4-93 */
static double CellTime(FS *s) {
  int ix, iy, iz;
  double tx, ty, tz, t;
  if (s->vx > 0) { ix = DOWN; tx = (M->lx[s->icell + 1] - s->x) / s->vx; }
  else if (s->vx < 0) { ix = UP; tx = -(s->x - M->lx[s->icell]) / s->vx; }
  else { ix = UP; tx = 1e99; }
  if (s->vy > 0) { iy = RIGHT; ty = (M->ly[s->jcell + 1] - s->y) / s->vy; }
  else if (s->vy < 0) { iy = LEFT; ty = -(s->y - M->ly[s->jcell]) / s->vy; }
  else { iy = RIGHT; ty = 1e99; }
  if (s->vz > 0) { iz = BACK; tz = (M->lz[s->kcell + 1] - s->z) / s->vz; }
  else if (s->vz < 0) { iz = FRONT; tz = -(s->z - M->lz[s->kcell]) / s->vz; }
  else { iz = FRONT; tz = 1e99; }
  s->idir = ix; t = tx;
  if (ty < t) { s->idir = iy; t = ty; }
  if (tz < t) { s->idir = iz; t = tz; }
  if (t < -MY_ZERO) t = 0;
  return t;
}

/* particle_fly, src/mcmodel.cpp:1874-2161 */
static void particle_fly(FS *s) {
  int getTet = 1, getCell = 1, getPh = 1, getImp = 1, getSurf = 1;
  double ssnl = 0, imprnl = 0, phrnl = 0;
  double TetTf = 0, CellTf = 0, PhScTf, ImpScTf, SurfScTf, Tf;
  s->Flag_Catch = 0;
  s->drop_reason = -1;
  s->SurfScGamma = 0;
  s->SurfRoughnessScRate = 0;
  s->SurfPhononScRate = 0;
  s->ImpScGamma = 0;
  s->Flag_SelfScatter = 0;
  s->idir = 0;
  s->itetdir = 0;
  load_cell(s);
  s->iband = B->ibt[s->itet];
  GetV(s);
  int loop = 0;
  while (s->left_time > 0) {
    GetV(s);
    loop++;
    s->t->n_events++;
    if (loop > LOOP_CAP) {
      s->Flag_Catch = 1;
      s->drop_reason = EMC_DROP_LOOPS;
      break;
    }
    if (getTet) {
      TetTf = tet_time_core(B, s->itet, s->isym, s->kx, s->ky, s->kz, s->Ex, s->Ey, s->Ez, &s->itetdir);
      getTet = 0;
    }
    if (getCell) {
      CellTf = CellTime(s);
      getCell = 0;
    }
    if (getPh) {
      phrnl = -orc_log(Random(s));
      getPh = 0;
    }
    PhScTf = phrnl / B->gamtet[s->itet];
    if (s->iband == B->bandof[PELEC]) {
      if (getImp) {
        imprnl = -orc_log(Random(s));
        getImp = 0;
      }
      s->ImpScGamma = GetImpScGamma(s);
      ImpScTf = imprnl / s->ImpScGamma;
    } else {
      ImpScTf = 2 * M->dt;
      s->ImpScGamma = 1.0 / M->scrt0;
    }
    const size_t c = CELL(s->icell, s->jcell, s->kcell);
    if (M->flag_surface && M->insurf[c]) {
      if (getSurf) {
        ssnl = -orc_log(Random(s));
        getSurf = 0;
      }
      if (s->par_type == PELEC) {
        s->SurfRoughnessScRate = s->f->e_rough[c];
        s->SurfPhononScRate = s->f->e_sph[c];
      } else {
        s->SurfRoughnessScRate = s->f->h_rough[c];
        s->SurfPhononScRate = s->f->h_sph[c];
      }
      s->SurfScGamma = s->SurfRoughnessScRate + s->SurfPhononScRate;
      if (s->SurfScGamma > 0 && ssnl > 0) SurfScTf = ssnl / s->SurfScGamma;
      else SurfScTf = 2 * M->dt;
    } else
      SurfScTf = 2 * M->dt;
    /* Min(6 args, &flag): first argument wins ties */
    int flag = 1;
    Tf = TetTf;
    if (Tf > CellTf) { Tf = CellTf; flag = 2; }
    if (Tf > PhScTf) { Tf = PhScTf; flag = 3; }
    if (Tf > ImpScTf) { Tf = ImpScTf; flag = 4; }
    if (Tf > SurfScTf) { Tf = SurfScTf; flag = 5; }
    if (Tf > s->left_time) { Tf = s->left_time; flag = 6; }
    if (Tf < MY_ZERO) {
      s->Flag_Catch = 1;
      s->drop_reason = EMC_DROP_TF_ZERO;
      break;
    }
    TetTf -= Tf; CellTf -= Tf; PhScTf -= Tf; ImpScTf -= Tf; SurfScTf -= Tf;
    s->left_time -= Tf;
    phrnl -= B->gamtet[s->itet] * Tf;
    imprnl -= s->ImpScGamma * Tf;
    ssnl -= s->SurfScGamma * Tf;
    s->kx += s->Ex * Tf;
    s->ky += s->Ey * Tf;
    s->kz += s->Ez * Tf;
    s->energy += (s->vx * s->Ex + s->vy * s->Ey + s->vz * s->Ez) * Tf;
    s->x += s->vx * Tf;
    s->y += s->vy * Tf;
    s->z += s->vz * Tf;
    if (s->energy < 0) {
      s->Flag_Catch = 1;
      s->drop_reason = EMC_DROP_NEG_ENERGY;
      break;
    }
    switch (flag) {
      case 1:
        HitTet(s);
        getTet = 1; getCell = 1;
        break;
      case 2:
        if (HitCell(s)) { getTet = 1; getCell = 1; }
        else {
          s->Flag_Catch = 1;
          s->drop_reason = EMC_DROP_OFF_MESH;
        }
        break;
      case 3:
        PhononScatter(s);
        if (!s->Flag_SelfScatter) {
          getTet = 1; getCell = 1;
          s->t->n_phonon++;
        }
        getPh = 1;
        break;
      case 4:
        if (s->par_type == PELEC) ElectronImpurityScatter(s);
        if (!s->Flag_SelfScatter) {
          getTet = 1; getCell = 1;
          s->t->n_impurity++;
        }
        getImp = 1;
        break;
      case 5:
        /* ParticleSurfaceScatter, src/ParSurfaceScatter.cpp:48-61 */
        if (Random(s) * s->SurfScGamma < s->SurfRoughnessScRate) {
          s->idir = M->surface_dir[M->nearest_surf[c]];
          Diffuse(s);
        } else {
          s->Flag_SelfScatter = 0;
          select_process(s, Random(s) * CALSCATTSUM(B, s->energy, s->iband));
        }
        if (!s->Flag_SelfScatter) { getTet = 1; getCell = 1; }
        getSurf = 1;
        break;
      default:
        break;
    }
    if (s->Flag_Catch) break;
  }
  if (s->Flag_Catch) {
    s->t->n_catch++;
    if (s->drop_reason >= 0) s->t->drops[s->drop_reason]++;
  }
}

static void fs_setup(FS *s, const emc_band_view *b, const emc_mesh_view *m, const emc_cell_fields *f,
                     uint64_t step, int flag_heat, emc_fly_tallies *t, int64_t *he, int64_t *hh, int64_t *rc,
                     int64_t *lc) {
  memset(s, 0, sizeof(*s));
  s->b = b; s->m = m; s->f = f; s->step = step; s->flag_heat = flag_heat; s->t = t;
  s->heat_e = he; s->heat_h = hh; s->rcur = rc; s->lcur = lc;
}

int orc_fly_step(const emc_band_view *b, const emc_mesh_view *m, const emc_cell_fields *f, emc_particles *p,
                 int64_t n, int64_t cap, uint64_t step, int flag_heat, emc_fly_tallies *t, int64_t *heat_e,
                 int64_t *heat_h, int64_t *rcur, int64_t *lcur, int64_t *n_out) {
  FS st, *s = &st;
  fs_setup(s, b, m, f, step, flag_heat, t, heat_e, heat_h, rcur, lcur);
  s->clone_cap = cap > n ? cap - n : 0;
  s->clones = (Clone *)malloc(sizeof(Clone) * (s->clone_cap > 0 ? s->clone_cap : 1));
  if (!s->clones) return -1;
  s->nclones = 0;
  for (int64_t q = 0; q < n; ++q) {
    if (p->i[q] < 0) continue;
    s->x = p->x[q]; s->y = p->y[q]; s->z = p->z[q];
    s->kx = p->kx[q]; s->ky = p->ky[q]; s->kz = p->kz[q];
    s->energy = p->energy[q]; s->charge = p->charge[q];
    s->icell = p->i[q]; s->jcell = p->j[q]; s->kcell = p->k[q];
    s->itet = p->itet[q]; s->isym = p->isym[q]; s->par_type = p->ptype[q];
    s->key = p->key[q];
    s->left_time = m->dt;
    s->nclone = 0;
    rng_init(&s->rng, s->key, step);
    particle_fly(s);
    if (s->Flag_Catch) {
      p->i[q] = -1;
      continue;
    }
    p->x[q] = s->x; p->y[q] = s->y; p->z[q] = s->z;
    p->kx[q] = s->kx; p->ky[q] = s->ky; p->kz[q] = s->kz;
    p->energy[q] = s->energy;
    p->i[q] = s->icell; p->j[q] = s->jcell; p->k[q] = s->kcell;
    p->itet[q] = s->itet; p->isym[q] = s->isym;
  }
  /* clones fly the rest of their step, in creation order; their own clones follow */
  int64_t done = 0, passes = 0;
  while (done < s->nclones) {
    int64_t end = s->nclones;
    passes++;
    for (; done < end; ++done) {
      Clone c = s->clones[done];
      s->x = c.x; s->y = c.y; s->z = c.z; s->kx = c.kx; s->ky = c.ky; s->kz = c.kz;
      s->energy = c.energy; s->charge = c.charge; s->left_time = c.left_time;
      s->icell = c.i; s->jcell = c.j; s->kcell = c.k; s->itet = c.itet; s->isym = c.isym;
      s->par_type = c.ptype; s->key = c.key;
      s->nclone = 0;
      rng_init(&s->rng, s->key, step);
      particle_fly(s);
      Clone *o = &s->clones[done];
      if (s->Flag_Catch) {
        o->i = -1;
        continue;
      }
      o->x = s->x; o->y = s->y; o->z = s->z; o->kx = s->kx; o->ky = s->ky; o->kz = s->kz;
      o->energy = s->energy; o->i = s->icell; o->j = s->jcell; o->k = s->kcell;
      o->itet = s->itet; o->isym = s->isym;
    }
  }
  t->n_clone_passes += passes;
  int64_t w = n;
  for (int64_t q = 0; q < s->nclones; ++q) {
    Clone *c = &s->clones[q];
    if (c->i < 0) continue;
    p->x[w] = c->x; p->y[w] = c->y; p->z[w] = c->z; p->kx[w] = c->kx; p->ky[w] = c->ky; p->kz[w] = c->kz;
    p->energy[w] = c->energy; p->charge[w] = c->charge;
    p->i[w] = c->i; p->j[w] = c->j; p->k[w] = c->k; p->itet[w] = c->itet; p->isym[w] = c->isym;
    p->ptype[w] = c->ptype; p->key[w] = c->key;
    ++w;
  }
  *n_out = w;
  free(s->clones);
  return 0;
}

/* ---- probes for tests */
int orc_get_state(const emc_band_view *b, uint64_t key, uint64_t step, double energy, int iband, int isym,
                  double k[3], int *itet, int *ndraws) {
  emc_fly_tallies t;
  memset(&t, 0, sizeof(t));
  FS st, *s = &st;
  memset(s, 0, sizeof(*s));
  s->b = b; s->t = &t;
  rng_init(&s->rng, key, step);
  s->energy = energy; s->iband = iband; s->isym = isym; s->itet = *itet;
  get_state(s);
  k[0] = s->kx; k[1] = s->ky; k[2] = s->kz;
  *itet = s->itet;
  *ndraws = (int)(s->rng.block * 2 - (s->rng.have ? 1 : 0));
  return s->Flag_SelfScatter;
}

/* Test hook: one EleImpScUpdateK (src/ElectronImpurityScatter.cpp:188-306) on an
 * electron at wedge-frame k with symmetry isym, energy `energy` and cell charge
 * density rho, draws from the stream of (key, step).  Returns 1 when the draw is
 * accepted (k, isym, itet, energy take the new state) and 0 on self-scattering. */
int orc_imp_update_k(const emc_band_view *b, const emc_mesh_view *m, uint64_t key, uint64_t step, double k[3],
                     int *isym, int *itet, double *energy, double rho) {
  emc_fly_tallies t;
  memset(&t, 0, sizeof(t));
  FS st, *s = &st;
  memset(s, 0, sizeof(*s));
  s->b = b; s->m = m; s->t = &t;
  rng_init(&s->rng, key, step);
  s->kx = k[0]; s->ky = k[1]; s->kz = k[2];
  s->isym = *isym; s->itet = *itet; s->energy = *energy; s->Rho = rho;
  s->par_type = PELEC;
  s->iband = B->ibt[*itet];
  EleImpScUpdateK(s);
  if (s->Flag_SelfScatter || s->Flag_Catch) return 0;
  k[0] = s->kx; k[1] = s->ky; k[2] = s->kz;
  *isym = s->isym; *itet = s->itet; *energy = s->energy;
  return 1;
}

int orc_getisym(const emc_band_view *b, const double k[3], double energy, int iband, int itet_hint, int *isym,
                int *itet) {
  emc_fly_tallies t;
  memset(&t, 0, sizeof(t));
  FS st, *s = &st;
  memset(s, 0, sizeof(*s));
  s->b = b; s->t = &t;
  s->kx = k[0]; s->ky = k[1]; s->kz = k[2];
  s->energy = energy; s->iband = iband; s->itet = itet_hint;
  Getisym(s);
  *isym = s->isym;
  *itet = s->itet;
  return s->Flag_Catch ? -1 : 0;
}

/* Test hook: the final energy electron phonon process `process` (5, 6, 9-12)
 * draws from the dispersive tables for energy eet, with the one random draw taken
 * from the stream of (key, step). */
double orc_fek(const emc_band_view *b, int process, double eet, uint64_t key, uint64_t step) {
  FS st;
  memset(&st, 0, sizeof st);
  FS *s = &st;
  s->b = b;
  rng_init(&s->rng, key, step);
  switch (process) {
    case 5: return FEKLOAB(s, eet);
    case 6: return FEKLOEM(s, eet);
    case 9: return FEKLAAB(s, eet);
    case 10: return FEKLAEM(s, eet);
    case 11: return FEKTOAB(s, eet);
    case 12: return FEKTOEM(s, eet);
    default: return eet;
  }
}
