// Per-particle full-band flight: update_particle / particle_fly and every
// callee (src/mcmodel.cpp:1874-2161, src/TetHit.cpp, src/CellHit.cpp,
// src/ParticleUtils.cpp:48-748, src/ElectronPhononScatter.cpp:5-497,
// src/HolePhononScatter.cpp, src/ElectronImpurityScatter.cpp,
// src/ParSurfaceScatter.cpp, src/heat.cpp:169-196).
//
// Written for one lane per particle: the state lives in registers, every table
// read is a gather from HBM/L2/MALL, tallies are int64 atomics.  The same code
// compiles for the host (initial sampling only uses get_state there).
//
// Semantics the reference leaves undefined or serial-order dependent, fixed here:
//  * Random(): per particle Philox stream, counter (draw block, step); see devmath.h.
//  * surface-phonon drand = Random()*CALSCATTSUM(E, band) (src/ParSurfaceScatter.cpp:72,124
//    read an uninitialised drand).
//  * SurfScGamma starts at 0 for each flight (a carried member in the reference).
//  * sign(0) = +1; pow(x,2) = x*x; pow(x,-2) = 1/(x*x).
//  * GENERATE clones fly the rest of the step in which they were made, with key
//    clone_key(parent, step, ordinal); HitCell across the outer mesh face drops
//    the particle (reason OFF_MESH) instead of re-flying it on a neighbour rank.
//  * the event loop stops after 65536 iterations (reason LOOPS).
#pragma once
#include <stdint.h>

#include "devmath.h"
#include "emc_const.h"
#include "tables.h"

namespace emc {

constexpr int FLY_LOOP_CAP = 65536;

// true if the condition holds for any active lane of the wave (device) / the
// condition itself (host)
#if defined(__HIP_DEVICE_COMPILE__)
#define EMC_ANY(p) (__ballot(p) != 0ull)
#else
#define EMC_ANY(p) (p)
#endif

EMC_HD inline void tally_add(unsigned long long* p, long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd(p, (unsigned long long)v);
#else
  *p += (unsigned long long)v;
#endif
}

EMC_HD inline long long fxd(double v, double scale) { return (long long)rint(v * scale); }

struct Flight {
  const BandDev& B;
  const MeshDev& M;
  const FieldsDev& F;
  const TallyDev& T;
  CloneRec* clones;
  unsigned long long* nclones;
  unsigned long long clone_cap;

  // particle
  double x, y, z, kx, ky, kz, energy, charge, left_time;
  int icell, jcell, kcell, itet, isym, par_type, iband;
  uint64_t key, step;
  // flight
  double vx, vy, vz, Ex, Ey, Ez, Rho, DA;
  double ImpScGamma, SurfScGamma, SurfRoughnessScRate, SurfPhononScRate;
  int itetdir, idir, Flag_Catch, Flag_SelfScatter, drop_reason, nclone, flag_heat;
  long long n_events, n_phonon, n_impurity;
  Rng rng;

  EMC_HD Flight(const BandDev& b, const MeshDev& m, const FieldsDev& f, const TallyDev& t)
      : B(b), M(m), F(f), T(t), clones(nullptr), nclones(nullptr), clone_cap(0) {}

  EMC_HD EMC_INL double Random() { return rng.next(); }
  EMC_HD EMC_INL int ms(int r, int s) const { return B.matsym[r * 48 + s]; }
  EMC_HD EMC_INL int indmat(int sx, int sy, int sz, int p0, int p1, int p2) const {
    return B.indmat[((((sx + 1) * 3 + (sy + 1)) * 3 + (sz + 1)) * 3 + p0) * 9 + p1 * 3 + p2];
  }
  EMC_HD EMC_INL size_t cell(int i, int j, int k) const { return ((size_t)j * M.cnx + i) * M.cnz + k; }
  EMC_HD EMC_INL size_t node(int i, int j, int k) const { return ((size_t)j * M.pnx + i) * M.pnz + k; }
  EMC_HD EMC_INL int corner(int c, int it) const { return B.tet[(size_t)c * B.nt + it]; }
  EMC_HD EMC_INL void count(int slot) { tally_add(&T.counters[slot], 1); }

  // in_wedge / OutWedge (src/ParticleUtils.cpp:293-305, src/Band.cpp:555-570)
  EMC_HD EMC_INL void in_wedge(double& xo, double& yo, double& zo, double xi, double yi, double zi) const {
    double v[3];
    v[ms(0, isym) - 1] = xi * ms(3, isym);
    v[ms(1, isym) - 1] = yi * ms(4, isym);
    v[ms(2, isym) - 1] = zi * ms(5, isym);
    xo = v[0]; yo = v[1]; zo = v[2];
  }
  EMC_HD EMC_INL void out_wedge(double& xo, double& yo, double& zo, double xi, double yi, double zi) const {
    const double v[3] = {xi, yi, zi};
    xo = v[ms(0, isym) - 1] * ms(3, isym);
    yo = v[ms(1, isym) - 1] * ms(4, isym);
    zo = v[ms(2, isym) - 1] * ms(5, isym);
  }
  EMC_HD EMC_INL void GetV() {
    const TetRec& r = B.tr[itet];
    out_wedge(vx, vy, vz, r.vt[0], r.vt[1], r.vt[2]);
  }
  EMC_HD EMC_INL void GetEnergy() {
    double dx, dy, dz;
    in_wedge(dx, dy, dz, kx, ky, kz);
    const CornerRec& q = B.cr[itet];
    dx -= q.x[0]; dy -= q.y[0]; dz -= q.z[0];
    const TetRec& r = B.tr[itet];
    energy = r.vt[3] + r.vt[0] * dx + r.vt[1] * dy + r.vt[2] * dz;
  }

  // ---- Band interpolators and areas (src/Band.cpp:194-352, 665-703, 1999-2321)
  EMC_HD EMC_INL static double fii(double x1, double y1, double z1, double x2, double y2, double z2, double x3,
                           double y3, double z3) {
    const double sbx = (y2 - y1) * (z3 - z1) - (y3 - y1) * (z2 - z1);
    const double sby = (x3 - x1) * (z2 - z1) - (x2 - x1) * (z3 - z1);
    const double sbz = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    return 0.50 * sqrt(sbx * sbx + sby * sby + sbz * sbz);
  }
  EMC_HD EMC_INL double surf(double e, int it) const {
    const CornerRec& q = B.cr[it];
    const int c0 = 0, c1 = 1, c2 = 2, c3 = 3;
    const double *X = q.x, *Y = q.y, *Z = q.z;
    const double e0 = q.e[0], e1 = q.e[1], e2 = q.e[2], e3 = q.e[3];
    double r;
    if (!(e > e0 && e < e3)) return 0.0;
    if (e <= e1) {
      r = (e - e0) / (e1 - e0);
      const double ax = r * (X[c1] - X[c0]), ay = r * (Y[c1] - Y[c0]), az = r * (Z[c1] - Z[c0]);
      r = (e - e0) / (e2 - e0);
      const double bx = r * (X[c2] - X[c0]), by = r * (Y[c2] - Y[c0]), bz = r * (Z[c2] - Z[c0]);
      r = (e - e0) / (e3 - e0);
      const double cx = r * (X[c3] - X[c0]), cy = r * (Y[c3] - Y[c0]), cz = r * (Z[c3] - Z[c0]);
      return fii(ax, ay, az, bx, by, bz, cx, cy, cz);
    }
    if (e > e1 && e < e2) {
      r = (e - e0) / (e2 - e0);
      const double ax = r * (X[c2] - X[c0]) + X[c0], ay = r * (Y[c2] - Y[c0]) + Y[c0], az = r * (Z[c2] - Z[c0]) + Z[c0];
      r = (e - e1) / (e2 - e1);
      const double bx = r * (X[c2] - X[c1]) + X[c1], by = r * (Y[c2] - Y[c1]) + Y[c1], bz = r * (Z[c2] - Z[c1]) + Z[c1];
      r = (e - e0) / (e3 - e0);
      const double cx = r * (X[c3] - X[c0]) + X[c0], cy = r * (Y[c3] - Y[c0]) + Y[c0], cz = r * (Z[c3] - Z[c0]) + Z[c0];
      r = (e - e1) / (e3 - e1);
      const double dx = r * (X[c3] - X[c1]) + X[c1], dy = r * (Y[c3] - Y[c1]) + Y[c1], dz = r * (Z[c3] - Z[c1]) + Z[c1];
      const double a1 = fii(ax, ay, az, bx, by, bz, cx, cy, cz);
      const double a2 = fii(dx, dy, dz, bx, by, bz, cx, cy, cz);
      const double a3 = fii(ax, ay, az, dx, dy, dz, cx, cy, cz);
      const double a4 = fii(ax, ay, az, bx, by, bz, dx, dy, dz);
      return 0.50 * (a1 + a2 + a3 + a4);
    }
    if (e > e2 && e < e3) {
      r = (e - e3) / (e0 - e3);
      const double ax = r * (X[c0] - X[c3]), ay = r * (Y[c0] - Y[c3]), az = r * (Z[c0] - Z[c3]);
      r = (e - e3) / (e1 - e3);
      const double bx = r * (X[c1] - X[c3]), by = r * (Y[c1] - Y[c3]), bz = r * (Z[c1] - Z[c3]);
      r = (e - e3) / (e2 - e3);
      const double cx = r * (X[c2] - X[c3]), cy = r * (Y[c2] - Y[c3]), cz = r * (Z[c2] - Z[c3]);
      return fii(ax, ay, az, bx, by, bz, cx, cy, cz);
    }
    return 0.0;
  }
  EMC_HD EMC_INL double fastsurf(double e, int it) { return fastsurf_r(e, it, B.sr[it]); }
  EMC_HD EMC_INL double fastsurf_r(double e, int it, const SurfRec& s) {
    double r = fastsurf_signed(e, it, s);
    if (r < 0.0) {
      if (T.counters) count(CNT_DROP0 + 6);
      r = 0.0;
    }
    return r;
  }
  // FASTSURF before the clamp of a negative value (which fastsurf_r tallies)
  EMC_HD EMC_INL double fastsurf_signed(double e, int it, const SurfRec& s) {
    const double e0 = s.e[0], e1 = s.e[1], e2 = s.e[2], e3 = s.e[3];
    double r = 0.0;
    if (e > e0 && e < e3) {
      if (e <= e1)
        r = s.faclow * (e - e0) * (e - e0);
      else if (e < e2) {
        if (e1 - e0 > 1e-10)
          r = s.faclow * (e - e0) * (e - e0) - s.facmed * (e - e1) * (e - e1);
        else
          r = surf(e, it);
      } else
        r = s.fachigh * (e - e3) * (e - e3);
    }
    return r;
  }
  EMC_HD EMC_INL double calscattsum(double e, int ib) const {
    if (e <= 0.0) return B.sumscatt[ib];
    if (e >= B.emax) return B.sumscatt[(size_t)MTAB * MNB + ib];
    const int itab = (int)(e / B.dtable);
    const double intp = (e - B.energy[itab]) / (B.energy[itab + 1] - B.energy[itab]);
    return B.sumscatt[(size_t)itab * MNB + ib] +
           intp * (B.sumscatt[(size_t)(itab + 1) * MNB + ib] - B.sumscatt[(size_t)itab * MNB + ib]);
  }
  // Band::CALDOS / CALDOSSUM (src/Band.cpp:2234-2295)
  EMC_HD EMC_INL double caldos(double e, int ib) const {
    const double* d = &B.dos[(size_t)ib * NTAB1];
    if (e <= 0e0) return d[0];
    if (e >= B.emax) return d[MTAB];
    const int itab = (int)(e / B.dtable);
    const double intp = (e - B.energy[itab]) / (B.energy[itab + 1] - B.energy[itab]);
    return d[itab] + intp * (d[itab + 1] - d[itab]);
  }
  EMC_HD EMC_INL double caldossum(double e, int ip) const {
    if (e <= 0e0) return B.sumdos[0 * 3 + ip];
    if (e >= B.emax) return B.sumdos[(size_t)MTAB * 3 + ip];
    const int itab = (int)(e / B.dtable);
    const double intp = (e - B.energy[itab]) / (B.energy[itab + 1] - B.energy[itab]);
    return B.sumdos[(size_t)itab * 3 + ip] +
           intp * (B.sumdos[(size_t)(itab + 1) * 3 + ip] - B.sumdos[(size_t)itab * 3 + ip]);
  }
  // select_kstate (src/mcmodel.cpp:3006-3057) for a new carrier of type par_type:
  // a Boltzmann energy accepted against the total DOS, a band by its share of the
  // DOS, a random symmetry operation, then get_state; tmp_const = 1 - exp(-emax).
  EMC_HD EMC_INL void sample_kstate(double tmp_const) {
    const int b0 = par_type == 0 ? B.bandofE : B.bandofH;
    const int nb = par_type == 0 ? B.nbandE : B.nbandH;
    itet = 0;
    // Every lane of the wave leaves this loop together (lanes that are done wait
    // without drawing), as in get_state's rejection loop: with lanes leaving on
    // different iterations and the random-number state live, the GPU build gave
    // other draws than the host build for most of the population (gfx950), after
    // changes elsewhere in the flight code.
    bool more = true;
    while (EMC_ANY(more)) {
      if (!more) continue;
      energy = -dm_log(1 - Random() * tmp_const);
      const double ddos = caldossum(energy, par_type);
      double rdos = Random() * B.dosmax[par_type];
      if (rdos > ddos) continue;
      int ib = b0;
      for (; ib < b0 + nb; ++ib) {
        rdos -= caldos(energy, ib);
        if (rdos < 0) break;
      }
      if (ib >= b0 + nb) ib = b0 + nb - 1;
      iband = ib;
      isym = ((int)(Random() * 48)) % 48;
      get_state();
      if (Flag_SelfScatter) continue;
      more = false;
    }
  }
  EMC_HD EMC_INL double caldostetmax(double e, int ib) const {
    const double* d = &B.dostetmax[(size_t)ib * NTAB1];
    if (e <= 0.0) return d[0];
    if (e >= B.emax) return d[MTAB];
    const int itab = (int)(e / B.dtable);
    return d[itab] > d[itab + 1] ? d[itab] : d[itab + 1];
  }
  EMC_HD EMC_INL void confbz(double& a, double& b, double& c) const {
    const double q = B.a0pi;
    for (int guard = 0; guard < 1024; ++guard) {  // a finite k folds in a few passes
      const bool out = fabs(a) > q || fabs(b) > q || fabs(c) > q || fabs(a) + fabs(b) + fabs(c) > 1.50 * q;
      if (!out) return;
      a = a > 0.0 ? a - q : a + q;
      b = b > 0.0 ? b - q : b + q;
      c = c > 0.0 ? c - q : c + q;
    }
  }

  // ---- tetrahedron search (src/ParticleUtils.cpp:331-625)
  EMC_HD EMC_INL bool contains(int it, double a, double b, double c) const {
    const TetRec& r = B.tr[it];
    const double d1 = r.n[0][0] * a + r.n[0][1] * b + r.n[0][2] * c - r.n[0][3];
    const double d2 = r.n[1][0] * a + r.n[1][1] * b + r.n[1][2] * c - r.n[1][3];
    const double d3 = r.n[2][0] * a + r.n[2][1] * b + r.n[2][2] * c - r.n[2][3];
    const double d4 = r.n[3][0] * a + r.n[3][1] * b + r.n[3][2] * c - r.n[3][3];
    return (d1 >= -MY_ZERO) && (d2 >= -MY_ZERO) && (d3 >= -MY_ZERO) && (d4 >= -MY_ZERO);
  }
  EMC_HD EMC_INL static bool in_sphere(const Sphere& s, double a, double b, double c) {
    const double dx = a - s.cx, dy = b - s.cy, dz = c - s.cz;
    return dx * dx + dy * dy + dz * dz <= s.r2;
  }
  // first tetrahedron of list[p0 .. p0+nn-1] (in list order) that holds the point;
  // the sphere array runs alongside the list, so most candidates cost 32 bytes
  // read in order instead of a dependent load of their record
  EMC_HD EMC_INL bool scan(const int* list, const Sphere* sph, int p0, int nn, double a, double b, double c) {
    const int end = p0 + nn;
    int itl = p0;
    // eight sphere tests per round, their loads independent of each other
    for (; itl + 8 <= end; itl += 8) {
      bool h[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) h[u] = in_sphere(sph[itl + u], a, b, c);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (h[u]) {
          const int it = list[itl + u];
          if (iband == B.tr[it].ibt && contains(it, a, b, c)) { itet = it; return true; }
        }
    }
    for (; itl < end; ++itl) {
      if (!in_sphere(sph[itl], a, b, c)) continue;
      const int it = list[itl];
      if (iband == B.tr[it].ibt && contains(it, a, b, c)) { itet = it; return true; }
    }
    return false;
  }
  // The answer of Getisym's list scans from the band's point-location octree: the
  // leaf holding (a, b, c) lists, in descending tetrahedron index, every tetrahedron
  // of the band that can hold the point, with its bucket ranges.  The first holding
  // entry in a bucket is what that bucket's scan returns (buckets run in descending
  // index), the last holding entry what the scan over all tetrahedra returns; the
  // buckets are tried in the scans' order: tlists bin, tlist bins ie, ie+1, ie-1.
  EMC_HD EMC_INL bool locate(double a, double b, double c) {
    const TetIdxRoot& R = B.troot[iband];
    if (!(a >= R.lo[0] && a <= R.hi[0] && b >= R.lo[1] && b <= R.hi[1] && c >= R.lo[2] && c <= R.hi[2]))
      return false;  // outside every tetrahedron of the band
    double l0 = R.lo[0], l1 = R.lo[1], l2 = R.lo[2], h0 = R.hi[0], h1 = R.hi[1], h2 = R.hi[2];
    TetIdxNode nd = B.tnode[R.node];
    while (nd.count < 0) {
      const double m0 = 0.5 * (l0 + h0), m1 = 0.5 * (l1 + h1), m2 = 0.5 * (l2 + h2);
      int ch = 0;
      if (a >= m0) { ch |= 1; l0 = m0; } else h0 = m0;
      if (b >= m1) { ch |= 2; l1 = m1; } else h1 = m1;
      if (c >= m2) { ch |= 4; l2 = m2; } else h2 = m2;
      nd = B.tnode[nd.first + ch];
    }
    int ies = (int)(energy * B.dlists);
    if (ies < 0) ies = 0;
    const bool use_s = ies < MWLES;
    int ie = (int)(energy * B.dlist);
    if (ie < 0) ie = 0;
    if (ie >= MWLE) ie = MWLE - 1;
    int bs = -1, b0 = -1, bup = -1, bdn = -1, ball = -1;
    for (int q = nd.first; q < nd.first + nd.count; ++q) {
      const TetIdxEntry& E = B.tent[q];
      if (!(a >= (double)E.lo[0] && a <= (double)E.hi[0] && b >= (double)E.lo[1] && b <= (double)E.hi[1] &&
            c >= (double)E.lo[2] && c <= (double)E.hi[2]))
        continue;
      if (!contains(E.t, a, b, c)) continue;
      const int lo = (int)(E.bins & 511u), hi = (int)((E.bins >> 9) & 511u);
      const int slo = (int)((E.bins >> 18) & 63u), shi = (int)((E.bins >> 24) & 63u);
      if (use_s && bs < 0 && slo <= ies && ies <= shi) bs = E.t;
      if (b0 < 0 && lo <= ie && ie <= hi) b0 = E.t;
      if (bup < 0 && ie + 1 < MWLE && lo <= ie + 1 && ie + 1 <= hi) bup = E.t;
      if (bdn < 0 && ie - 1 >= 0 && lo <= ie - 1 && ie - 1 <= hi) bdn = E.t;
      ball = E.t;
      if (bs >= 0 || (!use_s && b0 >= 0)) break;  // nothing later can take precedence
    }
    const int r = bs >= 0 ? bs : b0 >= 0 ? b0 : bup >= 0 ? bup : bdn >= 0 ? bdn : ball;
    if (r < 0) return false;
    itet = r;
    return true;
  }
  EMC_HD EMC_INL void Getisym() {
    int mink, maxk;
    double tmin, tmax;
    if (fabs(kx) > fabs(ky)) { tmin = fabs(ky); tmax = fabs(kx); mink = 1; maxk = 0; }
    else { tmin = fabs(kx); tmax = fabs(ky); mink = 0; maxk = 1; }
    if (fabs(kz) < tmin) mink = 2;
    else if (fabs(kz) > tmax) maxk = 2;
    // position of each axis in the sorted |k| order: 0 largest, 1 middle, 2 smallest
    const int p0 = maxk == 0 ? 0 : (mink == 0 ? 2 : 1);
    const int p1 = maxk == 1 ? 0 : (mink == 1 ? 2 : 1);
    const int p2 = maxk == 2 ? 0 : (mink == 2 ? 2 : 1);
    isym = indmat(kx < 0 ? -1 : 1, ky < 0 ? -1 : 1, kz < 0 ? -1 : 1, p0, p1, p2);
    double a, b, c;
    in_wedge(a, b, c, kx, ky, kz);
    if (iband == B.tr[itet].ibt && contains(itet, a, b, c)) return;
    if (B.tnode) {
      if (!locate(a, b, c)) {
        Flag_Catch = 1;
        drop_reason = EMC_DROP_NO_TET_;
      }
      return;
    }
    int ie = (int)(energy * B.dlists);
    if (ie < 0) ie = 0;
    if (ie < MWLES && scan(B.tlists, B.tlists_s, B.ptlists[ie * MNB + iband], B.ntlists[ie * MNB + iband], a, b, c)) return;
    ie = (int)(energy * B.dlist);
    if (ie < 0) ie = 0;
    if (ie >= MWLE) ie = MWLE - 1;
    if (scan(B.tlist, B.tlist_s, B.ptlist[ie * MNB + iband], B.ntlist[ie * MNB + iband], a, b, c)) return;
    if (ie + 1 < MWLE && scan(B.tlist, B.tlist_s, B.ptlist[(ie + 1) * MNB + iband], B.ntlist[(ie + 1) * MNB + iband], a, b, c))
      return;
    if (ie - 1 >= 0 && scan(B.tlist, B.tlist_s, B.ptlist[(ie - 1) * MNB + iband], B.ntlist[(ie - 1) * MNB + iband], a, b, c))
      return;
    for (int it = 0; it < B.nt; ++it)
      if (in_sphere(B.tsph[it], a, b, c) && iband == B.tr[it].ibt && contains(it, a, b, c)) { itet = it; return; }
    Flag_Catch = 1;
    drop_reason = EMC_DROP_NO_TET_;
  }
  static constexpr int EMC_DROP_NO_TET_ = 2;
  EMC_HD EMC_INL void GetisymBZ() {
    double a, b, c;
    in_wedge(a, b, c, kx, ky, kz);
    if (iband == B.tr[itet].ibt && contains(itet, a, b, c)) return;
    int ie = (int)(energy * B.dlist);
    if (ie < 0) ie = 0;
    if (ie >= MWLE) ie = MWLE - 1;
    if (scan(B.slist, B.slist_s, B.pslist[ie * MNB + iband], B.nslist[ie * MNB + iband], a, b, c)) return;
    Getisym();
  }

  // ---- final state on the iso-energy surface (src/ParticleUtils.cpp:48-284)
  EMC_HD EMC_INL void getK() {
    const CornerRec& q = B.cr[itet];
    const double *X = q.x, *Y = q.y, *Z = q.z, *E = q.e;
    const int c0 = 0, c1 = 1, c2 = 2, c3 = 3;
    const double rv1 = 1.0 - sqrt(Random());
    const double rv2 = Random() * (1.0 - rv1);
    if ((energy < E[c0]) || (energy > E[c3])) energy = E[c0] + (E[c3] - E[c0]) * Random();
    const double e = energy;
    double r, xkl, ykl, zkl;
    if (e <= E[c1]) {
      r = (e - E[c0]) / (E[c1] - E[c0]);
      const double ax = r * (X[c1] - X[c0]), ay = r * (Y[c1] - Y[c0]), az = r * (Z[c1] - Z[c0]);
      r = (e - E[c0]) / (E[c2] - E[c0]);
      const double bx = r * (X[c2] - X[c0]), by = r * (Y[c2] - Y[c0]), bz = r * (Z[c2] - Z[c0]);
      r = (e - E[c0]) / (E[c3] - E[c0]);
      const double cx = r * (X[c3] - X[c0]), cy = r * (Y[c3] - Y[c0]), cz = r * (Z[c3] - Z[c0]);
      xkl = rv1 * (bx - ax) + rv2 * (cx - ax) + ax + X[c0];
      ykl = rv1 * (by - ay) + rv2 * (cy - ay) + ay + Y[c0];
      zkl = rv1 * (bz - az) + rv2 * (cz - az) + az + Z[c0];
    } else if (e < E[c2]) {
      r = (e - E[c0]) / (E[c2] - E[c0]);
      const double ax = r * (X[c2] - X[c0]) + X[c0], ay = r * (Y[c2] - Y[c0]) + Y[c0], az = r * (Z[c2] - Z[c0]) + Z[c0];
      r = (e - E[c1]) / (E[c2] - E[c1]);
      const double bx = r * (X[c2] - X[c1]) + X[c1], by = r * (Y[c2] - Y[c1]) + Y[c1], bz = r * (Z[c2] - Z[c1]) + Z[c1];
      r = (e - E[c0]) / (E[c3] - E[c0]);
      const double cx = r * (X[c3] - X[c0]) + X[c0], cy = r * (Y[c3] - Y[c0]) + Y[c0], cz = r * (Z[c3] - Z[c0]) + Z[c0];
      r = (e - E[c1]) / (E[c3] - E[c1]);
      const double dx = r * (X[c3] - X[c1]) + X[c1], dy = r * (Y[c3] - Y[c1]) + Y[c1], dz = r * (Z[c3] - Z[c1]) + Z[c1];
      const double a1 = fii(ax, ay, az, bx, by, bz, cx, cy, cz);
      const double a2 = fii(dx, dy, dz, bx, by, bz, cx, cy, cz);
      const double a3 = fii(ax, ay, az, dx, dy, dz, cx, cy, cz);
      const double a4 = fii(ax, ay, az, bx, by, bz, dx, dy, dz);
      const double ar = Random() * (a1 + a2 + a3 + a4);
      double px, py, pz, qx, qy, qz, ox, oy, oz;  // triangle (origin o, edges p-o, q-o)
      if (ar < a1) { ox = ax; oy = ay; oz = az; px = bx; py = by; pz = bz; qx = cx; qy = cy; qz = cz; }
      else if (ar < (a1 + a2)) { ox = dx; oy = dy; oz = dz; px = bx; py = by; pz = bz; qx = cx; qy = cy; qz = cz; }
      else if (ar < (a1 + a2 + a3)) { ox = ax; oy = ay; oz = az; px = dx; py = dy; pz = dz; qx = cx; qy = cy; qz = cz; }
      else { ox = ax; oy = ay; oz = az; px = bx; py = by; pz = bz; qx = dx; qy = dy; qz = dz; }
      xkl = rv1 * (px - ox) + rv2 * (qx - ox) + ox;
      ykl = rv1 * (py - oy) + rv2 * (qy - oy) + oy;
      zkl = rv1 * (pz - oz) + rv2 * (qz - oz) + oz;
    } else {
      r = (e - E[c3]) / (E[c0] - E[c3]);
      const double ax = r * (X[c0] - X[c3]), ay = r * (Y[c0] - Y[c3]), az = r * (Z[c0] - Z[c3]);
      r = (e - E[c3]) / (E[c1] - E[c3]);
      const double bx = r * (X[c1] - X[c3]), by = r * (Y[c1] - Y[c3]), bz = r * (Z[c1] - Z[c3]);
      r = (e - E[c3]) / (E[c2] - E[c3]);
      const double cx = r * (X[c2] - X[c3]), cy = r * (Y[c2] - Y[c3]), cz = r * (Z[c2] - Z[c3]);
      xkl = rv1 * (bx - ax) + rv2 * (cx - ax) + ax + X[c3];
      ykl = rv1 * (by - ay) + rv2 * (cy - ay) + ay + Y[c3];
      zkl = rv1 * (bz - az) + rv2 * (cz - az) + az + Z[c3];
    }
    out_wedge(kx, ky, kz, xkl, ykl, zkl);
  }

  EMC_HD EMC_INL void get_state() {
    const double e = energy;
    if (e <= 0.0) { Flag_SelfScatter = 1; return; }
    const double maxdostet = caldostetmax(e, iband);
    int itab = (int)(e * B.dlists);
    const bool smallfl = itab < MWLES;
    if (!smallfl) {
      itab = (int)(e * B.dlist);
      if (itab >= MWLE) itab = MWLE - 1;
    }
    const int nted = smallfl ? B.ntlists[itab * MNB + iband] : B.ntlist[itab * MNB + iband];
    const int base = smallfl ? B.ptlists[itab * MNB + iband] : B.ptlist[itab * MNB + iband];
    const int* list = smallfl ? B.tlists : B.tlist;
    if (nted == 0) { Flag_SelfScatter = 1; return; }
    int count2 = 0;
    // Rejection loop.  On the GPU every lane of the wave leaves it together
    // (lanes that are done wait without drawing): a loop whose lanes exit on
    // different iterations, with the random-number state live after it, gave
    // wrong draws in getK for some lanes on gfx950.  Draws per particle are
    // the same as in the plain loop.
    bool again = true;
    while (EMC_ANY(again)) {
      if (!again) continue;
      itet = list[(int)(nted * Random()) + base];
      count2++;
      const SurfRec& sr = B.sr[itet];
      if ((count2 < 10000) && !((sr.e[0] < e) && (sr.e[3] > e))) continue;
      const double dostet = fastsurf_r(e, itet, sr) / sr.vgt;
      if (((Random() * maxdostet) > dostet) && (count2 < 10000)) continue;
      again = false;
    }
    if (count2 >= 10000) Flag_SelfScatter = 1;
    else { Flag_SelfScatter = 0; getK(); }
  }

  // get_state() in pieces, for its wave-parallel form in the slow flight kernel
  // (get_state_wave) and for get_state_chunked(), the same search on one thread.
  // Tries that run long are taken GS_CHUNK draws at a time: draw n of a particle
  // depends only on its key, the step and n (Philox is counter based), so every
  // position of a chunk can be formed at once and evaluated both as the first
  // draw of a try (the tetrahedron it picks, whether that spans the energy, its
  // DOS weight) and as the acceptance draw of a try that started one position
  // earlier; gs_walk() then follows the loop's control flow over those bits.  The
  // tries, the draws consumed, the tetrahedron and the tallies are get_state()'s.
  static constexpr int GS_CAP = 10000;  // get_state()'s bound on the tries
  static constexpr int GS_CHUNK = 128;  // draws per chunk (two per lane of a wave)
  // get_state() up to its loop: false when no draw is made (self-scattering set)
  EMC_HD EMC_INL bool gs_begin(double& maxdostet, int& nted, int& base, bool& small) {
    const double e = energy;
    if (e <= 0.0) { Flag_SelfScatter = 1; return false; }
    maxdostet = caldostetmax(e, iband);
    int itab = (int)(e * B.dlists);
    small = itab < MWLES;
    if (!small) {
      itab = (int)(e * B.dlist);
      if (itab >= MWLE) itab = MWLE - 1;
    }
    nted = small ? B.ntlists[itab * MNB + iband] : B.ntlist[itab * MNB + iband];
    base = small ? B.ptlists[itab * MNB + iband] : B.ptlist[itab * MNB + iband];
    if (nted == 0) { Flag_SelfScatter = 1; return false; }
    return true;
  }
  // one try of the loop: true when the search goes on
  EMC_HD EMC_INL bool gs_try(double e, double maxdostet, int nted, int base, bool small, int& count2) {
    const int* list = small ? B.tlists : B.tlist;
    itet = list[(int)(nted * Random()) + base];
    count2++;
    const SurfRec& sr = B.sr[itet];
    if ((count2 < GS_CAP) && !((sr.e[0] < e) && (sr.e[3] > e))) return true;
    const double dostet = fastsurf_r(e, itet, sr) / sr.vgt;
    if (((Random() * maxdostet) > dostet) && (count2 < GS_CAP)) return true;
    return false;
  }
  // get_state() after its loop
  EMC_HD EMC_INL void gs_end(int count2) {
    if (count2 >= GS_CAP) Flag_SelfScatter = 1;
    else { Flag_SelfScatter = 0; getK(); }
  }
  // draw u taken as the first draw of a try: the tetrahedron, whether it spans
  // e, its DOS weight when it does and whether FASTSURF was negative there
  EMC_HD EMC_INL void gs_eval(double e, int nted, int base, bool small, double u, int& it, bool& span, bool& neg,
                              double& dostet) {
    const int* list = small ? B.tlists : B.tlist;
    it = list[(int)(nted * u) + base];
    const SurfRec& sr = B.sr[it];
    span = (sr.e[0] < e) && (sr.e[3] > e);
    neg = false;
    dostet = 0.0;
    if (span) {
      double r = fastsurf_signed(e, it, sr);
      if (r < 0.0) { neg = true; r = 0.0; }
      dostet = r / sr.vgt;
    }
  }
  // The loop over one chunk.  Bit q of span / acc / neg (q < 64 in word 0, else
  // word 1): the try whose first draw is position q picks a tetrahedron spanning
  // the energy / accepts it with the draw at q + 1 / has a negative FASTSURF.
  // count2 tries were made before the chunk.  A try that needs the draw after
  // the chunk is left to the next one.
  struct GsWalk {
    int consumed;  // draws used
    int count2;    // tries made in all
    int pos;       // first draw of the last try, when done
    int nneg;      // FASTSURF clamps among the tries
    bool done;
  };
  // The two words of each set are passed apart (no array indexed at run time).
  // The tries are found with word operations instead of one by one: a try that
  // starts at a spanning position takes the next position as its acceptance
  // draw, so within a run of spanning positions the tries start at every other
  // position from the run's first, and a run begins with a try.  Runs are split
  // by the parity of their first position with a carry (adding the run's first
  // bit clears the run); the tries of the chunk are then counted with popcounts
  // and the try the limit falls on is selected.
  static EMC_HD EMC_INL uint64_t gs_span_starts(uint64_t S, uint64_t first_odd) {
    const uint64_t E = 0x5555555555555555ull;  // even positions
    const uint64_t RS = S & ~(S << 1);         // first positions of runs
    const uint64_t even_first = RS & E & ~first_odd;
    const uint64_t run_even = S & ~(S + even_first);
    return (run_even & E) | (S & ~run_even & ~E);
  }
  // set bits at positions 0..q of the 128-bit set (m0, m1)
  static EMC_HD EMC_INL int gs_pop_upto(uint64_t m0, uint64_t m1, int q) {
    if (q < 64) return __builtin_popcountll(m0 & ((2ull << q) - 1ull));
    return __builtin_popcountll(m0) + __builtin_popcountll(m1 & ((2ull << (q - 64)) - 1ull));
  }
  // position of the k-th set bit (k >= 1), or -1
  static EMC_HD EMC_INL int gs_select(uint64_t m0, uint64_t m1, int k) {
    uint64_t m = m0;
    int off = 0;
    const int p0 = __builtin_popcountll(m0);
    if (k > p0) { k -= p0; m = m1; off = 64; }
    if (k > __builtin_popcountll(m)) return -1;
    for (int i = 1; i < k; ++i) m &= m - 1ull;
    return off + __builtin_ctzll(m);
  }
  static EMC_HD EMC_INL GsWalk gs_walk(uint64_t span0, uint64_t span1, uint64_t acc0, uint64_t acc1, uint64_t neg0,
                                       uint64_t neg1, int count2) {
    // spanning tries (TS) and all tries (T); position 64 is an acceptance draw
    // when a spanning try starts at 63
    const uint64_t ts0 = gs_span_starts(span0, 0);
    const uint64_t a64 = ts0 >> 63;
    const uint64_t ts1 = gs_span_starts(span1, a64 & span1 & 1ull);
    const uint64_t t0 = ~(ts0 << 1), t1 = ~((ts1 << 1) | a64);
    const uint64_t ac0 = acc0 & ts0, ac1 = acc1 & ts1, ng0 = neg0 & ts0, ng1 = neg1 & ts1;
    const int qa = ac0 ? __builtin_ctzll(ac0) : ac1 ? 64 + __builtin_ctzll(ac1) : -1;  // first accepted
    const int tc = gs_select(t0, t1, GS_CAP - count2);  // the try the limit falls on
    const int last = (tc >= 0 && (qa < 0 || tc <= qa)) ? tc : qa;
    GsWalk w{GS_CHUNK, count2, -1, 0, false};
    if (last >= 0 && last + 1 < GS_CHUNK) {
      w.consumed = last + 2;
      w.count2 = count2 + gs_pop_upto(t0, t1, last);
      w.nneg = gs_pop_upto(ng0, ng1, last);
      w.pos = last;
      w.done = true;
    } else if (last >= 0 || (ts1 >> 63)) {
      // the chunk's last try needs the draw after it: left to the next chunk
      w.consumed = GS_CHUNK - 1;
      w.count2 = count2 + gs_pop_upto(t0, t1, GS_CHUNK - 2);
      w.nneg = gs_pop_upto(ng0, ng1, GS_CHUNK - 2);
    } else {
      w.count2 = count2 + __builtin_popcountll(t0) + __builtin_popcountll(t1);
      w.nneg = __builtin_popcountll(ng0) + __builtin_popcountll(ng1);
    }
    return w;
  }
  // Draw position q of a chunk that starts at Philox block blk, after the spare
  // draw sp when have is set; block words as in Rng::next.
  static EMC_HD EMC_INL double gs_word(const uint32_t w[4], int part) {
    const uint64_t v = part ? ((((uint64_t)w[3] << 32) | w[2]) >> 12) : ((((uint64_t)w[1] << 32) | w[0]) >> 12);
    return ((double)v + 0.5) * 2.220446049250313080847263336181640625e-16;
  }
  // the stream state after k draws of a chunk: block advanced past every block
  // whose first word was used; have set when the next draw is a second word
  // (the caller then sets spare to it, unless have was set and k == 0)
  static EMC_HD EMC_INL void gs_advance(uint32_t& blk, int& have, int k) {
    blk += (uint32_t)(have ? k / 2 : (k + 1) / 2);
    have = have ? !(k & 1) : (k & 1);
  }
  EMC_HD double gs_draw(uint32_t blk, int have, double sp, int q) const {
    const int r = q - have;
    if (r < 0) return sp;
    uint32_t w[4] = {blk + (uint32_t)(r >> 1), rng.c1, rng.c2, 0u};
    philox4x32_10(w, rng.k0, rng.k1);
    return gs_word(w, r & 1);
  }
  // get_state() with the tries after the first `serial` taken in chunks, on one
  // thread: the restatement of get_state_wave that the host checks
  EMC_HD void get_state_chunked(int serial) {
    const double e = energy;
    double maxdostet = 0;
    int nted = 0, base = 0;
    bool small = false;
    if (!gs_begin(maxdostet, nted, base, small)) return;
    int count2 = 0;
    bool again = true;
    while (again && count2 < serial) again = gs_try(e, maxdostet, nted, base, small, count2);
    while (again) {
      uint64_t span[2] = {0, 0}, acc[2] = {0, 0}, neg[2] = {0, 0};
      int its[GS_CHUNK];
      double u[GS_CHUNK], dost[GS_CHUNK];
      for (int q = 0; q < GS_CHUNK; ++q) u[q] = gs_draw(rng.block, rng.have, rng.spare, q);
      for (int q = 0; q < GS_CHUNK; ++q) {
        bool sp = false, ng = false;
        gs_eval(e, nted, base, small, u[q], its[q], sp, ng, dost[q]);
        if (sp) span[q >> 6] |= 1ull << (q & 63);
        if (ng) neg[q >> 6] |= 1ull << (q & 63);
        if (sp && q + 1 < GS_CHUNK && !((u[q + 1] * maxdostet) > dost[q])) acc[q >> 6] |= 1ull << (q & 63);
      }
      const GsWalk w = gs_walk(span[0], span[1], acc[0], acc[1], neg[0], neg[1], count2);
      if (w.nneg && T.counters) tally_add(&T.counters[CNT_DROP0 + 6], w.nneg);
      count2 = w.count2;
      if (w.done) { itet = its[w.pos]; again = false; }
      uint32_t blk = rng.block;
      int have = rng.have;
      const double next = gs_draw(blk, have, rng.spare, w.consumed);
      const bool keep = have && w.consumed == 0;
      gs_advance(blk, have, w.consumed);
      if (have && !keep) rng.spare = next;
      rng.block = blk;
      rng.have = have;
    }
    gs_end(count2);
  }

  // TALIND (src/ParticleUtils.cpp:626-748)
  EMC_HD EMC_INL void TALIND(int ityp) {
    int ital = 1;
    const double ax = fabs(kx), ay = fabs(ky), az = fabs(kz);
    if ((ax >= ay) && (ax >= az)) {
      if (ityp == 3) { const int t = (int)(Random() * 4.0 + 1.0); ital = t <= 2 ? t + 2 : t + 2; }
      else ital = (kx >= 0.0) == (ityp == 1) ? 1 : 2;
    }
    if ((ay >= az) && (ay >= ax)) {
      if (ityp == 3) { const int t = (int)(Random() * 4.0 + 1.0); ital = t <= 2 ? t : t + 2; }
      else ital = (ky >= 0.0) == (ityp == 1) ? 3 : 4;
    }
    if ((az > ay) && (az > ax)) {
      if (ityp == 3) ital = (int)(Random() * 4.0 + 1.0);
      else ital = (kz >= 0.0) == (ityp == 1) ? 5 : 6;
    }
    isym = (int)(Random() * 8.0) + 8 * (ital - 1);
  }

  // heat_to_point (src/heat.cpp:169-196)
  EMC_HD EMC_INL void heat_to_point(double e, unsigned long long* heat) {
    if (!heat) return;
    const int i = icell, j = jcell, k = kcell;
    const double rx = (x - M.lx[i]) / M.dx[i];
    const double ry = (y - M.ly[j]) / M.dy[j];
    const double rz = (z - M.lz[k]) / M.dz[k];
    tally_add(&heat[node(i, j, k)], fxd(e * (1 - ry) * (1 - rx) * (1 - rz), EMC_FX_HEAT_));
    tally_add(&heat[node(i + 1, j, k)], fxd(e * (1 - ry) * rx * (1 - rz), EMC_FX_HEAT_));
    tally_add(&heat[node(i, j, k + 1)], fxd(e * (1 - ry) * (1 - rx) * rz, EMC_FX_HEAT_));
    tally_add(&heat[node(i + 1, j, k + 1)], fxd(e * (1 - ry) * rx * rz, EMC_FX_HEAT_));
    tally_add(&heat[node(i, j + 1, k)], fxd(e * ry * (1 - rx) * (1 - rz), EMC_FX_HEAT_));
    tally_add(&heat[node(i + 1, j + 1, k)], fxd(e * ry * rx * (1 - rz), EMC_FX_HEAT_));
    tally_add(&heat[node(i, j + 1, k + 1)], fxd(e * ry * (1 - rx) * rz, EMC_FX_HEAT_));
    tally_add(&heat[node(i + 1, j + 1, k + 1)], fxd(e * ry * rx * rz, EMC_FX_HEAT_));
  }
  static constexpr double EMC_FX_HEAT_ = 1099511627776.0;     // 2^40
  static constexpr double EMC_FX_CONTACT_ = 1099511627776.0;  // 2^40

  // ---- phonon scattering (src/ElectronPhononScatter.cpp:5-497, src/HolePhononScatter.cpp)
  struct Saved {
    double e, kx, ky, kz;
    int iband, itet, isym;
  };
  EMC_HD EMC_INL Saved save() const { return Saved{energy, kx, ky, kz, iband, itet, isym}; }
  EMC_HD EMC_INL void restore(const Saved& s) {
    energy = s.e; kx = s.kx; ky = s.ky; kz = s.kz; iband = s.iband; itet = s.itet; isym = s.isym;
  }
  // Final energy of electron phonon process p from the dispersive tables (runs
  // with the quantum correction, src/ElectronPhononScatter.cpp:148-163): p = 5/6
  // g-LO absorption/emission (FEKLOAB/FEKLOEM, :702-768, :632-700), 9/10 f-LA
  // (FEKLAAB/FEKLAEM, :770-889), 11/12 f-TO (FEKTOAB/FEKTOEM, :891-1022).  The
  // energy in units of the phonon's picks an energy bin (finer bins at low
  // energy); one draw walks that bin's distribution over final-energy classes,
  // and class i shifts the energy by the phonon's, widened or narrowed by
  // (i/2) steps of the bin's width (up for odd i, down for even i).
  EMC_HD EMC_INL double dispersive(int p, double eet) {
    const bool lo = p <= 6, em = (p & 1) == 0;
    const int tab = p == 6 ? 0 : p == 5 ? 1 : p == 12 ? 2 : p == 11 ? 3 : p == 10 ? 4 : 5;
    const double hw = lo ? B.templog : p <= 10 ? B.templaf : B.temptof;
    const double ie = eet / hw;
    int it, ei;
    double ed;
    if (em) {
      if (ie < 1.0) return p == 10 ? -1.0e-10 : 1.0e-10;
      if (ie < 3.05) { it = (int)((ie - 1.0) * 10.0 + 0.5); it = it < 1 ? 1 : it; ei = it; ed = lo ? 1.0 : 0.25; }
      else if (ie < 5.625) { it = (int)((ie - 3.0) * 4.0 + 0.5); it = it < 1 ? 1 : it; ei = 20 + it; ed = lo ? 2.0 : 0.5; }
      else if (ie < 13.25) { it = (int)((ie - 5.5) * 2.0 + 0.5); it = it < 1 ? 1 : it; ei = 30 + it; ed = lo ? 4.0 : 0.5; }
      else if (lo && ie < 28.5) { it = (int)(ie - 13.0 + 0.5); it = it < 1 ? 1 : it; ei = 45 + it; ed = 8.0; }
      else if (lo) { it = (int)((ie - 28.0) + 0.5); it = it < 1 ? 1 : it > 60 ? 60 : it; ei = 60 + it; ed = 12.0; }
      else { it = (int)((ie - 13.0) + 0.5); it = it < 1 ? 1 : it > 115 ? 115 : it; ei = 45 + it; ed = 1.0; }
    } else {
      if (ie < 1.05) { it = (int)(ie * 10.0 + 0.5); it = it < 1 ? 1 : it; ei = it; ed = lo ? 1.0 : 0.25; }
      else if (ie < 3.625) { it = (int)((ie - 1.0) * 4.0 + 0.5); it = it < 1 ? 1 : it; ei = 10 + it; ed = lo ? 2.0 : 0.25; }
      else if (ie < 11.25) { it = (int)((ie - 3.5) * 2.0 + 0.5); it = it < 1 ? 1 : it; ei = 20 + it; ed = lo ? 4.0 : 0.5; }
      else if (lo && ie < 26.5) { it = (int)((ie - 11.0) + 0.5); it = it < 1 ? 1 : it; ei = 35 + it; ed = 8.0; }
      else if (lo) { it = (int)((ie - 26.0) + 0.5); it = it < 1 ? 1 : it > 60 ? 60 : it; ei = 50 + it; ed = 12.0; }
      else { it = (int)((ie - 11.0) + 0.5); it = it < 1 ? 1 : it > 60 ? 115 : it; ei = 35 + it; ed = 1.0; }  // > 60 jumps to 115, as the reference
    }
    const int R = tab < 2 ? 201 : 51;  // PTAB_ROWS
    const double* col = B.ptab[tab] + (size_t)(ei - 1) * R;
    const double ran = Random();
    int i = 0;
    double ksum = 0.0;
    if (p == 6) {  // FEKLOEM wraps before it adds
      while (ksum < ran) {
        ++i;
        if (i == R) { ksum = 1.1; i = 1; }
        ksum += col[i - 1];
      }
    } else {
      while (ksum < ran) {
        ++i;
        ksum += col[i - 1];
        if (i == R) { ksum = 1.1; i = 1; }
      }
    }
    const int j = i % 2, k = i / 2;
    const double w = (lo ? 0.01 : 0.04) * ed * k;
    double ef = em ? (j == 1 ? eet - hw * (1.0 - w) : eet - hw * (1.0 + w))
                   : (j == 1 ? eet + hw * (1.0 + w) : eet + hw * (1.0 - w));
    if ((p == 9 || p == 10) && ef <= 0.0) ef = 1.0e-5 * hw;
    return ef;
  }

  // The scattering routines that end in get_state are split in two around it:
  // prep_* does what comes before get_state and returns which finish() applies
  // after it, so get_state (with getK) and the process selection are compiled
  // once in the event loop (a smaller flight kernel).  The order of operations
  // and random draws is that of the routines of the reference.
  enum { POST_NONE = 0, POST_EPSCAT = 1, POST_SCATII = 2, POST_HPSCAT = 3, POST_DIFFUSE = 4 };

  // EPSCAT (src/ElectronPhononScatter.cpp:98-195), ESCATII / HSCATII (:406-497,
  // src/HolePhononScatter.cpp) and HPSCAT (src/HolePhononScatter.cpp), up to get_state
  EMC_HD EMC_INL int prep_process(int iscat, int np, bool elec, Saved& s0) {
    s0 = save();
    if (((iscat % np) + 1) == np) {  // impact ionisation
      iband = iscat / np + (par_type == PELEC ? B.bandofE : B.bandofH);
      energy = (energy - B.sieg) / 3.0;
      isym = ((int)(48 * Random()) % 48);
      return POST_SCATII;
    }
    if (elec) {
      iband = iscat / B.scpre + B.bandofE;
      const int p = iscat % B.scpre;
      if (p == 0) TALIND(1);
      else if (p <= 12) {
        if (B.ptab[0] && (p == 5 || p == 6 || p >= 9)) {
          energy = dispersive(p, energy);
        } else {
          const double temps[6] = {B.temptag, B.templag, B.templog, B.temptaf, B.templaf, B.temptof};
          const double hw = temps[(p - 1) >> 1];
          energy = (p & 1) ? energy + hw : energy - hw;
        }
        TALIND(p <= 6 ? 2 : 3);
      }
      return POST_EPSCAT;
    }
    iband = iscat / B.scprh + B.bandofH;
    const int p = iscat % B.scprh;
    if (p == 1) energy = energy + B.temphop;
    else if (p == 2) energy = energy - B.temphop;
    isym = ((int)(48 * Random()) % 48);
    return POST_HPSCAT;
  }
  // Diffuse (src/ParSurfaceScatter.cpp, src/CellHit.cpp), up to get_state
  EMC_HD EMC_INL int prep_diffuse(Saved& s0) {
    s0 = save();
    TALIND(1);
    return POST_DIFFUSE;
  }
  EMC_HD EMC_INL void finish(int post, const Saved& s0) {
    if (post == POST_EPSCAT || post == POST_HPSCAT) {
      if (Flag_SelfScatter) restore(s0);
      if (flag_heat) heat_to_point((energy - s0.e) * fabs(charge), post == POST_EPSCAT ? T.heat_e : T.heat_h);
    } else if (post == POST_SCATII) {
      if (Flag_SelfScatter) restore(s0);
    } else if (post == POST_DIFFUSE) {
      if (Flag_SelfScatter) { iband = s0.iband; itet = s0.itet; isym = s0.isym; kx = s0.kx; ky = s0.ky; kz = s0.kz; }
      GetV();
      double v = 0.0;
      int axis = 0;
      bool flip = false, known = true;
      switch (idir) {
        case UP: v = vx; axis = 0; flip = v < 0.0; break;
        case RIGHT: v = vy; axis = 1; flip = v > 0.0; break;
        case DOWN: v = vx; axis = 0; flip = v > 0.0; break;
        case LEFT: v = vy; axis = 1; flip = v < 0.0; break;
        case FRONT: v = vz; axis = 2; flip = v < 0.0; break;
        case BACK: v = vz; axis = 2; flip = v > 0.0; break;
        default: known = false; break;
      }
      if (known) {
        if (v == 0.0) { iband = s0.iband; itet = s0.itet; isym = s0.isym; kx = s0.kx; ky = s0.ky; kz = s0.kz; }
        if (flip) mirror(axis);
      }
    }
  }
  // rate of process iscat (CALSCATTE / CALSCATTH element, src/Band.cpp:2098-2231)
  EMC_HD EMC_INL double process_rate(bool elec, int iscat, int itab, double intp, int where) const {
    const int np = elec ? B.scpre : B.scprh;
    const int is = iscat % np, f = iscat / np;
    const int ini = iband - (elec ? B.bandofE : B.bandofH);
    const int nb = elec ? 4 : 3;
    const double* dos = elec ? B.dose : B.dosh;
    const double* sc = elec ? B.scatte : B.scatth;
    const double* ii = elec ? B.scattiie : B.scattiih;
    const double* drow;
    if (is < np - 1) {
      const double k = sc[((size_t)is * nb + f) * nb + ini];
      drow = &dos[((size_t)is * nb + f) * NTAB1];
      if (where < 0) return k * drow[0];
      if (where > 0) return k * drow[MTAB];
      return k * (drow[itab] + intp * (drow[itab + 1] - drow[itab]));
    }
    drow = &ii[((size_t)f * nb + ini) * NTAB1];
    if (where < 0) return drow[0];
    if (where > 0) return drow[MTAB];
    return drow[itab] + intp * (drow[itab + 1] - drow[itab]);
  }
  // process selection of the phonon scatterings (src/ElectronPhononScatter.cpp:5-95,
  // src/HolePhononScatter.cpp:4-60): cumulative search over the rates
  EMC_HD EMC_INL int select_prep(double drand, Saved& s0) {
    const bool elec = par_type == PELEC;
    const int np = elec ? B.scpre : B.scprh;
    const int nmax = np * (elec ? B.nbandE : B.nbandH);
    const double e = energy;
    int where = 0, itab = 0;
    double intp = 0.0;
    if (e <= 0.0) where = -1;
    else if (e >= B.emax) where = 1;
    else {
      itab = (int)(e / B.dtable);
      intp = (e - B.energy[itab]) / (B.energy[itab + 1] - B.energy[itab]);
    }
    int iscat = -1;
    double dscat = 0.0;
    while (drand > dscat) {
      iscat = iscat + 1;
      if (iscat >= nmax) { Flag_SelfScatter = 1; break; }
      dscat = dscat + process_rate(elec, iscat, itab, intp, where);
    }
    if (Flag_SelfScatter) return POST_NONE;
    return prep_process(iscat, np, elec, s0);
  }
  // ElectronPhononScatter / HolePhononScatter: self-scattering test against gamtet
  EMC_HD EMC_INL int prep_phonon(Saved& s0) {
    const double scasum = calscattsum(energy, iband);
    const double g = B.tr[itet].gam;
    const double drand = scasum > g ? Random() * scasum : Random() * g;
    if (scasum < drand) { Flag_SelfScatter = 1; return POST_NONE; }
    Flag_SelfScatter = 0;
    return select_prep(drand, s0);
  }

  // ---- ionised impurities (src/ElectronImpurityScatter.cpp:6-306)
  EMC_HD EMC_INL double GetImpScRate(double eel) const {
    if ((eel * M.eV0 < 0.120) && (DA * M.conc0 >= 1e22)) {
      const double fr = M.frickel;
      const double eebeta = fr * 4 * PI * Rho / (2.0 * B.meld * M.eps_si);
      if (M.flag_surface && M.flag_simp && M.insurf[cell(icell, jcell, kcell)]) return 0.0;
      const double fr2 = (fr * Rho) * (fr * Rho);
      return DA * B.meld * sqrt(2.0 * B.meld * eel) / (PI * fr2 * (1.0 + 4.0 * eel * B.melt / (eebeta * B.meld)));
    }
    return 0.0;
  }
  EMC_HD EMC_INL double GetImpScGamma() const {
    const double e0 = B.tr[itet].vt[3], e3 = B.tr[itet].e3;
    const double eemax = M.frickel * 4 * PI * Rho / (8.0 * B.melt * M.eps_si);
    double gammax;
    if (eemax < e0) gammax = GetImpScRate(e0);
    else if (eemax > e3) gammax = GetImpScRate(e3);
    else gammax = GetImpScRate(eemax);
    const double lo = 1.0 / M.scrt0;
    return gammax > lo ? gammax : lo;
  }
  EMC_HD EMC_INL void EleImpScUpdateK() {
    double xkl, ykl, zkl;
    in_wedge(xkl, ykl, zkl, kx, ky, kz);
    xkl = xkl - 0.85 * B.a0pi;
    xkl = xkl * sqrt(B.meld / B.mell);
    ykl = ykl * sqrt(B.meld / B.melt);
    zkl = zkl * sqrt(B.meld / B.melt);
    const double betaq = M.frickel * 4 * PI * Rho / M.eps_si;
    const double eebeta = betaq / (2.0 * B.meld);
    const double alfa = eebeta * B.meld / (2.0 * energy * B.melt);
    double pr = Random() / (1.0 + 0.5 * alfa);
    const double costr = 1.0 - alfa * pr / (1.0 - pr);
    const double sintr = sqrt(1.0 - costr * costr);
    pr = 2.0 * PI * Random();
    const double cospr = dm_cos(pr), sinpr = dm_sin(pr);
    const double ca = sqrt(xkl * xkl + ykl * ykl);
    const double cb = sqrt(ca * ca + zkl * zkl);
    const double xkll = cb * sintr * cospr, ykll = cb * sintr * sinpr, zkll = cb * costr;
    const double cosphi = xkl / ca, sinphi = ykl / ca, costheta = zkl / cb, sintheta = ca / cb;
    const double xklll = cosphi * costheta * xkll - sinphi * ykll + cosphi * sintheta * zkll;
    const double yklll = sinphi * costheta * xkll + cosphi * ykll + sinphi * sintheta * zkll;
    const double zklll = -sintheta * xkll + costheta * zkll;
    const double tn = betaq + B.mell / B.meld * (xkl - xklll) * (xkl - xklll) +
                      B.melt / B.meld * ((ykl - yklll) * (ykl - yklll) + (zkl - zklll) * (zkl - zklll));
    const double tm = betaq + B.melt / B.meld * (xkl - xklll) * (xkl - xklll) +
                      B.melt / B.meld * ((ykl - yklll) * (ykl - yklll) + (zkl - zklll) * (zkl - zklll));
    const double dn = 1.0 / (tn * tn), dnmax = 1.0 / (tm * tm);
    if (dn > Random() * dnmax) {
      xkl = xklll * sqrt(B.mell / B.meld);
      ykl = yklll * sqrt(B.melt / B.meld);
      zkl = zklll * sqrt(B.melt / B.meld);
      xkl = xkl + 0.85 * B.a0pi;
      out_wedge(kx, ky, kz, xkl, ykl, zkl);
      confbz(kx, ky, kz);
      Getisym();
      GetEnergy();
      Flag_SelfScatter = 0;
    } else
      Flag_SelfScatter = 1;
  }
  EMC_HD EMC_INL void ElectronImpurityScatter() {
    const double gamimp = GetImpScRate(energy);
    if (gamimp < Random() * ImpScGamma) Flag_SelfScatter = 1;
    else EleImpScUpdateK();
  }

  // ---- boundaries (src/CellHit.cpp)
  EMC_HD EMC_INL void mirror(int axis) {
    int a = ms(3, isym), b = ms(4, isym), c = ms(5, isym);
    if (axis == 0) { kx = -kx; a = -a; }
    else if (axis == 1) { ky = -ky; b = -b; }
    else { kz = -kz; c = -c; }
    isym = indmat(a, b, c, ms(0, isym) - 1, ms(1, isym) - 1, ms(2, isym) - 1);
  }
  EMC_HD EMC_INL void Reflect() {
    if (idir == UP || idir == DOWN) mirror(0);
    else if (idir == RIGHT || idir == LEFT) mirror(1);
    else if (idir == FRONT || idir == BACK) mirror(2);
  }
  EMC_HD EMC_INL bool neighbour(int& ii, int& jj, int& kk) const {
    ii = icell; jj = jcell; kk = kcell;
    switch (idir) {
      case UP: --ii; break;
      case DOWN: ++ii; break;
      case LEFT: --jj; break;
      case RIGHT: ++jj; break;
      case FRONT: --kk; break;
      case BACK: ++kk; break;
    }
    return ii >= 0 && ii < M.cnx && jj >= 0 && jj < M.cny && kk >= 0 && kk < M.cnz;
  }
  EMC_HD EMC_INL void contact_tally(int icont, int col, long long v) {
    if (icont < 0 || icont >= EMC_MAX_CONTACT_ || icont >= M.ncontact) { count(CNT_DROP0 + 7); return; }
    tally_add(&T.contact[icont * 5 + col], v);
  }
  static constexpr int EMC_MAX_CONTACT_ = 16;
  EMC_HD EMC_INL void Generate() {
    count(CNT_GEN);
    const int icont = M.attached[cell(icell, jcell, kcell)] - 1;
    if (icont < 0) count(CNT_DROP0 + 7);
    else {
      contact_tally(icont, 0, 1);
      contact_tally(icont, 1, fxd(charge, EMC_FX_CONTACT_));
      contact_tally(icont, 2, fxd(energy, EMC_FX_CONTACT_));
    }
    if (idir == RIGHT) tally_add(&T.rcur[jcell + 1], fxd(charge, EMC_FX_CONTACT_));
    if (idir == LEFT) tally_add(&T.lcur[jcell], fxd(charge, EMC_FX_CONTACT_));
    int ii, jj, kk;
    neighbour(ii, jj, kk);
    const int ord = nclone++;
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned long long slot = atomicAdd(nclones, 1ull);
#else
    const unsigned long long slot = (*nclones)++;
#endif
    if (slot >= clone_cap) { count(CNT_DROP0 + 8); return; }
    CloneRec& c = clones[slot];
    c.x = x; c.y = y; c.z = z; c.kx = kx; c.ky = ky; c.kz = kz;
    c.energy = energy; c.charge = charge; c.left_time = left_time;
    c.key = clone_key(key, step, ord);
    c.pk = pk_make(par_type, ii, jj, kk);
    c.itet = itet; c.isym = isym; c.pad = 0;
  }
  EMC_HD EMC_INL void Catch(bool gate) {
    int ii, jj, kk, icont = -1;
    if (neighbour(ii, jj, kk)) icont = M.attached[cell(ii, jj, kk)] - 1;
    if (icont < 0) count(CNT_DROP0 + 7);
    if (!gate) {
      if (idir == RIGHT) tally_add(&T.rcur[jcell + 1], fxd(charge, EMC_FX_CONTACT_));
      if (idir == LEFT) tally_add(&T.lcur[jcell], fxd(charge, EMC_FX_CONTACT_));
    }
    if (icont >= 0) {
      contact_tally(icont, 3, 1);
      contact_tally(icont, 4, fxd(charge, EMC_FX_CONTACT_));
      contact_tally(icont, 2, fxd(energy, EMC_FX_CONTACT_));  // the reference adds to EnergyGen
    }
  }
  EMC_HD EMC_INL void load_cell() {
    const size_t c = cell(icell, jcell, kcell);
    Rho = F.rho[c];
    DA = M.da[c];
    const double sg = M.charge_sign[par_type];
    if (par_type == 0) { Ex = sg * F.ex[c]; Ey = sg * F.ey[c]; Ez = sg * F.ez[c]; }
    else { Ex = sg * F.hx[c]; Ey = sg * F.hy[c]; Ez = sg * F.hz[c]; }
  }
  // a Diffuse at an oxide face is left to the caller (diffuse = true): it reads no
  // cell data, so running it after load_cell() changes nothing
  EMC_HD EMC_INL bool HitCell(bool& diffuse) {
    const int imot = M.motion[cell(icell, jcell, kcell) * 6 + idir];
    int ii, jj, kk;
    const bool inside = neighbour(ii, jj, kk);
    if ((imot == PASS || imot == GENERATE || imot == GENREF) && !inside) return false;
    if (imot == PASS) {
      switch (idir) {
        case UP: x = M.lx[icell]; icell--; break;
        case DOWN: x = M.lx[icell + 1]; icell++; break;
        case RIGHT:
          y = M.ly[jcell + 1];
          jcell++;
          tally_add(&T.rcur[jcell], fxd(charge, EMC_FX_CONTACT_));
          break;
        case LEFT:
          y = M.ly[jcell];
          tally_add(&T.lcur[jcell], fxd(charge, EMC_FX_CONTACT_));
          jcell--;
          break;
        case FRONT: z = M.lz[kcell]; kcell--; break;
        case BACK: z = M.lz[kcell + 1]; kcell++; break;
      }
    } else if (imot == REFLECT) {
      Reflect();
    } else if (imot == SCATTOX) {
      if (M.flag_surface) Reflect();
      else if (Random() > B.difpr[par_type]) Reflect();
      else diffuse = true;
    } else if (imot == PERIOD) {
      if (idir == UP) x = M.lx[icell + 1];
      else if (idir == DOWN) x = M.lx[icell];
      else if (idir == LEFT) y = M.ly[jcell + 1];
      else if (idir == RIGHT) y = M.ly[jcell];
      else if (idir == FRONT) z = M.lz[kcell + 1];
      else if (idir == BACK) z = M.lz[kcell];
    } else if (imot == GENERATE || imot == GENREF) {
      Generate();
      Reflect();
    } else if (imot == CATCH || imot == CATCHGATE) {
      Catch(imot == CATCHGATE);
      Flag_Catch = 1;
      drop_reason = -1;
    } else {
      Flag_Catch = 1;
      drop_reason = 5;
      return true;
    }
    load_cell();
    return true;
  }

  // TetTime (src/TetHit.cpp:4-106)
  EMC_HD EMC_INL double TetTime() {
    double xkl, ykl, zkl, xfl, yfl, zfl;
    in_wedge(xkl, ykl, zkl, kx, ky, kz);
    in_wedge(xfl, yfl, zfl, Ex, Ey, Ez);
    const TetRec& r = B.tr[itet];
    double best = 0.0;
    for (int f = 0; f < 4; ++f) {
      const double nf = r.n[f][0] * xfl + r.n[f][1] * yfl + r.n[f][2] * zfl;
      double tfp;
      if (nf < 0) {
        tfp = (r.n[f][3] - (r.n[f][0] * xkl + r.n[f][1] * ykl + r.n[f][2] * zkl)) / nf;
        if (tfp < 0) tfp = 0;
      } else
        tfp = 1e99;
      if (f == 0 || tfp < best) { best = tfp; itetdir = f; }
    }
    return best;
  }
  // HitTet (src/TetHit.cpp:109-142)
  EMC_HD EMC_INL void HitTet() {
    const int nb = B.tr[itet].neib[itetdir];
    if (nb >= 0) itet = nb;
    else if (nb == -5 || nb == -6) {
      const int ii = -nb - 2;
      kx = kx + B.xbz[ii * 48 + isym];
      ky = ky + B.ybz[ii * 48 + isym];
      kz = kz + B.zbz[ii * 48 + isym];
      isym = B.newisym[ii * 48 + isym];
      GetisymBZ();
      if (Flag_Catch) return;
      GetEnergy();
    } else
      isym = B.newisym[(-nb - 2) * 48 + isym];
  }
  // CellTime (src/CellHit.cpp:4-93)
  EMC_HD EMC_INL double CellTime() {
    int ix, iy, iz;
    double tx, ty, tz, t;
    if (vx > 0) { ix = DOWN; tx = (M.lx[icell + 1] - x) / vx; }
    else if (vx < 0) { ix = UP; tx = -(x - M.lx[icell]) / vx; }
    else { ix = UP; tx = 1e99; }
    if (vy > 0) { iy = RIGHT; ty = (M.ly[jcell + 1] - y) / vy; }
    else if (vy < 0) { iy = LEFT; ty = -(y - M.ly[jcell]) / vy; }
    else { iy = RIGHT; ty = 1e99; }
    if (vz > 0) { iz = BACK; tz = (M.lz[kcell + 1] - z) / vz; }
    else if (vz < 0) { iz = FRONT; tz = -(z - M.lz[kcell]) / vz; }
    else { iz = FRONT; tz = 1e99; }
    idir = ix; t = tx;
    if (ty < t) { idir = iy; t = ty; }
    if (tz < t) { idir = iz; t = tz; }
    if (t < -MY_ZERO) t = 0;
    return t;
  }

  // particle_fly (src/mcmodel.cpp:1874-2161), split into the parts the fast path reuses.
  struct Loop {
    bool getTet, getCell, getPh, getImp, getSurf;
    double ssnl, imprnl, phrnl, TetTf, CellTf;
    int loop;
  };
  EMC_HD EMC_INL void fly_begin(Loop& L) {
    L.getTet = L.getCell = L.getPh = L.getImp = L.getSurf = true;
    L.ssnl = L.imprnl = L.phrnl = 0;
    L.TetTf = L.CellTf = 0;
    L.loop = 0;
    Flag_Catch = 0; drop_reason = -1;
    SurfScGamma = 0; SurfRoughnessScRate = 0; SurfPhononScRate = 0; ImpScGamma = 0;
    Flag_SelfScatter = 0; idir = 0; itetdir = 0;
    load_cell();
    iband = B.tr[itet].ibt;
  }
  // candidate times of the current iteration; returns the Min(...) flag and Tf
  EMC_HD EMC_INL int next_event(Loop& L, double& Tf, double& gam, size_t& c) {
    if (L.getTet) { L.TetTf = TetTime(); L.getTet = false; }
    if (L.getCell) { L.CellTf = CellTime(); L.getCell = false; }
    if (L.getPh) { L.phrnl = -dm_log(Random()); L.getPh = false; }
    gam = B.tr[itet].gam;
    const double PhScTf = L.phrnl / gam;
    double ImpScTf;
    if (iband == B.bandofE) {
      if (L.getImp) { L.imprnl = -dm_log(Random()); L.getImp = false; }
      ImpScGamma = GetImpScGamma();
      ImpScTf = L.imprnl / ImpScGamma;
    } else {
      ImpScTf = 2 * M.dt;
      ImpScGamma = 1.0 / M.scrt0;
    }
    c = cell(icell, jcell, kcell);
    double SurfScTf;
    if (M.flag_surface && M.insurf[c]) {
      if (L.getSurf) { L.ssnl = -dm_log(Random()); L.getSurf = false; }
      if (par_type == PELEC) { SurfRoughnessScRate = F.e_rough[c]; SurfPhononScRate = F.e_sph[c]; }
      else { SurfRoughnessScRate = F.h_rough[c]; SurfPhononScRate = F.h_sph[c]; }
      SurfScGamma = SurfRoughnessScRate + SurfPhononScRate;
      SurfScTf = (SurfScGamma > 0 && L.ssnl > 0) ? L.ssnl / SurfScGamma : 2 * M.dt;
    } else
      SurfScTf = 2 * M.dt;
    // Min(6 args, &flag): the first argument wins ties (src/Constant_GlobalVariable_MathFunction.cpp:92-118)
    int flag = 1;
    Tf = L.TetTf;
    if (Tf > L.CellTf) { Tf = L.CellTf; flag = 2; }
    if (Tf > PhScTf) { Tf = PhScTf; flag = 3; }
    if (Tf > ImpScTf) { Tf = ImpScTf; flag = 4; }
    if (Tf > SurfScTf) { Tf = SurfScTf; flag = 5; }
    if (Tf > left_time) { Tf = left_time; flag = 6; }
    return flag;
  }
  // free flight for Tf (src/mcmodel.cpp:2009-2046)
  EMC_HD EMC_INL void advance(Loop& L, double Tf, double gam) {
    L.TetTf -= Tf; L.CellTf -= Tf;
    left_time -= Tf;
    L.phrnl -= gam * Tf;
    L.imprnl -= ImpScGamma * Tf;
    L.ssnl -= SurfScGamma * Tf;
    kx += Ex * Tf; ky += Ey * Tf; kz += Ez * Tf;
    energy += (vx * Ex + vy * Ey + vz * Ez) * Tf;
    x += vx * Tf; y += vy * Tf; z += vz * Tf;
  }

  // One iteration of fly()'s event loop, split at the final-state draw so that
  // the slow kernel can run get_state() with the whole wave (get_state_wave):
  // event_head() runs up to it (false: the particle stops, fly() leaves its
  // loop; post says whether a final state is drawn), get_state() follows when
  // post is set, then event_tail() (false: the particle was caught).
  EMC_HD EMC_INL bool event_head(Loop& L, int& flag, int& post, Saved& s0) {
    GetV();
    L.loop++;
    n_events++;
    if (L.loop > FLY_LOOP_CAP) { Flag_Catch = 1; drop_reason = 3; return false; }
    double Tf, gam;
    size_t c;
    flag = next_event(L, Tf, gam, c);
    if (Tf < MY_ZERO) { Flag_Catch = 1; drop_reason = 0; return false; }
    advance(L, Tf, gam);
    if (energy < 0) { Flag_Catch = 1; drop_reason = 1; return false; }
    post = POST_NONE;
    switch (flag) {
      case 1:
        HitTet();
        L.getTet = L.getCell = true;
        break;
      case 2: {
        bool diffuse = false;
        if (HitCell(diffuse)) L.getTet = L.getCell = true;
        else { Flag_Catch = 1; drop_reason = 4; }
        if (diffuse) post = prep_diffuse(s0);
        break;
      }
      case 3:
        post = prep_phonon(s0);
        break;
      case 4:
        if (par_type == PELEC) ElectronImpurityScatter();
        if (!Flag_SelfScatter) { L.getTet = L.getCell = true; n_impurity++; }
        L.getImp = true;
        break;
      case 5:
        // ParticleSurfaceScatter (src/ParSurfaceScatter.cpp:48-61)
        if (Random() * SurfScGamma < SurfRoughnessScRate) {
          idir = M.surface_dir[M.nearest_surf[c]];
          post = prep_diffuse(s0);
        } else {
          Flag_SelfScatter = 0;
          post = select_prep(Random() * calscattsum(energy, iband), s0);
        }
        break;
      default:
        break;
    }
    return true;
  }
  EMC_HD EMC_INL bool event_tail(Loop& L, int flag, int post, const Saved& s0) {
    if (post != POST_NONE) finish(post, s0);
    if (flag == 3) {
      if (!Flag_SelfScatter) { L.getTet = L.getCell = true; n_phonon++; }
      L.getPh = true;
    } else if (flag == 5) {
      if (!Flag_SelfScatter) L.getTet = L.getCell = true;
      L.getSurf = true;
    }
    return !Flag_Catch;
  }
  EMC_HD EMC_INL void fly() {
    Loop L;
    fly_begin(L);
    while (left_time > 0) {
      int flag = 0, post = POST_NONE;
      Saved s0{};
      if (!event_head(L, flag, post, s0)) break;
      if (post != POST_NONE) get_state();
      if (!event_tail(L, flag, post, s0)) break;
    }
  }

  // The common case of a step: up to FAST_EVENTS events that are tetrahedron
  // crossings inside the zone or wedge changes, cell crossings by PASS, REFLECT,
  // SCATTOX with surface scattering on, or PERIOD, then the end of the step.
  // fast_step() runs exactly the iterations of fly() for those events.  It bails
  // out when anything else comes first (scattering, a Brillouin-zone face,
  // contacts, leaving the mesh, a second current tally): the particle in memory is
  // untouched and fly() redoes it from the start with the same random numbers.
  // A y-plane current of a PASS crossing is held in (pend_kind, pend_j) and
  // added by commit_pending() only once the particle has finished here.
  static constexpr int FAST_EVENTS = 16;
  EMC_HD EMC_INL bool fast_cell_crossing(int& pend_kind, int& pend_j) {
    const int imot = M.motion[cell(icell, jcell, kcell) * 6 + idir];
    if (imot == PASS) {
      int ii, jj, kk;
      if (!neighbour(ii, jj, kk)) return false;
      switch (idir) {
        case UP: x = M.lx[icell]; icell--; break;
        case DOWN: x = M.lx[icell + 1]; icell++; break;
        case RIGHT:
          if (pend_kind) return false;
          y = M.ly[jcell + 1];
          jcell++;
          pend_kind = 1; pend_j = jcell;
          break;
        case LEFT:
          if (pend_kind) return false;
          y = M.ly[jcell];
          pend_kind = 2; pend_j = jcell;
          jcell--;
          break;
        case FRONT: z = M.lz[kcell]; kcell--; break;
        case BACK: z = M.lz[kcell + 1]; kcell++; break;
      }
    } else if (imot == REFLECT || (imot == SCATTOX && M.flag_surface)) {
      Reflect();
    } else if (imot == PERIOD) {
      if (idir == UP) x = M.lx[icell + 1];
      else if (idir == DOWN) x = M.lx[icell];
      else if (idir == LEFT) y = M.ly[jcell + 1];
      else if (idir == RIGHT) y = M.ly[jcell];
      else if (idir == FRONT) z = M.lz[kcell + 1];
      else if (idir == BACK) z = M.lz[kcell];
    } else
      return false;
    load_cell();
    return true;
  }
  // One iteration of fly()'s event loop on the fast path, after fly_begin() and
  // with pend_kind = pend_j = 0 at the start of the particle, entered while
  // left_time > 0: FAST_MORE (left_time still > 0), FAST_DONE (the step ended
  // here) or FAST_BAIL (fly() must redo the particle).
  enum { FAST_MORE = 0, FAST_DONE = 1, FAST_BAIL = 2 };
  EMC_HD EMC_INL int fast_step(Loop& L, int& pend_kind, int& pend_j) {
    GetV();
    L.loop++;
    n_events++;
    if (L.loop > FAST_EVENTS) return FAST_BAIL;
    double Tf, gam;
    size_t c;
    const int flag = next_event(L, Tf, gam, c);
    if (Tf < MY_ZERO || (flag >= 3 && flag <= 5)) return FAST_BAIL;
    advance(L, Tf, gam);
    if (energy < 0) return FAST_BAIL;
    if (flag == 1) {
      // HitTet without the Brillouin-zone fold
      const int nb = B.tr[itet].neib[itetdir];
      if (nb == -5 || nb == -6) return FAST_BAIL;
      if (nb >= 0) itet = nb;
      else isym = B.newisym[(-nb - 2) * 48 + isym];
      L.getTet = L.getCell = true;
    } else if (flag == 2) {
      if (!fast_cell_crossing(pend_kind, pend_j)) return FAST_BAIL;
      L.getTet = L.getCell = true;
    }
    return left_time > 0 ? FAST_MORE : FAST_DONE;
  }
  EMC_HD EMC_INL void commit_pending(int pend_kind, int pend_j) {
    if (pend_kind == 1) tally_add(&T.rcur[pend_j], fxd(charge, EMC_FX_CONTACT_));
    else if (pend_kind == 2) tally_add(&T.lcur[pend_j], fxd(charge, EMC_FX_CONTACT_));
  }
};

}  // namespace emc
