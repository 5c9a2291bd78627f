// Full-band tables: restatement of Band::IELEC and its callees (src/Band.cpp).
// Arithmetic order follows the reference so that table values match what the
// reference computes from the same files with the same libm.
#include "band.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace emc {

// ---------------------------------------------------------------- physics
// MeshQuantities::init_phpysical_parameter, src/mcmodel.cpp:4007-4122 (Si branch).
void phys_init(Phys& p, double T) {
  p.frickel = 1.0;
  p.psi_si = 4.05;
  double sia0 = 5.43e-10, sirho = 2.33e3, siul = 9.05e3, siut = 9.05e3;
  p.T0 = T;
  p.Tn = T / 300.0;
  p.eV0 = BOLTZ * p.T0;
  p.em0 = EM;
  p.hq0 = PLANCK;
  p.ec0 = EC;
  p.rmom0 = std::sqrt((p.em0 / p.ec0) * p.eV0);
  p.spr0 = p.hq0 / p.rmom0;
  p.spk0 = 1.0 / p.spr0;
  p.time0 = p.hq0 / p.eV0;
  p.velo0 = p.spr0 / p.time0;
  p.cvr = CLIGHT / p.velo0;
  p.pot0 = p.eV0;
  p.field0 = p.pot0 / p.spr0;
  p.conc0 = p.spk0 * p.spk0 * p.spk0;
  p.dens0 = p.em0 * p.conc0;
  p.dpc0 = p.field0;
  p.scrt0 = 1.0 / p.time0;
  p.curr0 = p.ec0 / p.time0;
  p.Nc = 2 * std::pow(1.08 * p.em0 * p.pot0 / (p.hq0 * p.hq0 * p.ec0 * 2 * PI), 1.5);
  p.N_cur = 2 * p.hq0 / (p.em0 * PI * PI) * std::pow((p.em0 * p.eV0) / (p.hq0 * p.hq0), 2) / p.ec0;
  p.sia0 = sia0 / p.spr0;
  p.sirho = sirho / p.dens0;
  p.siul = siul / p.velo0;
  p.siut = siut / p.velo0;
  if (p.T0 < 190.0)
    p.sieg = (1.170 + 1.059e-5 * p.T0 - 6.05e-7 * p.T0 * p.T0) / p.eV0;
  else if (p.T0 < 250.0)
    p.sieg = (1.17850 - 9.025e-5 * p.T0 - 3.05e-7 * p.T0 * p.T0) / p.eV0;
  else
    p.sieg = (1.2060 - 2.730e-4 * p.T0) / p.eV0;
  p.a0pi = TWOPI / p.sia0;
  const double FSC = EC / (PLANCK * CLIGHT * 4.0 * PI * EPS0);
  for (int i = 0; i < MNMaterialType; ++i) p.eps[i] = 0.0;
  p.eps[VACUUM] = 1.00 / (4 * PI * p.cvr * FSC);
  p.eps[OXIDE] = 3.90 / (4 * PI * p.cvr * FSC);
  p.eps[SILICON] = 11.70 / (4 * PI * p.cvr * FSC);
  p.SurfSc_ail = 2.0e-9 / p.spr0;
  p.SurfSc_delta = 0.1126755e-9 / p.spr0;
  p.SurfSc_XIph = 12.0;
  p.SurfSc_theta = 2.857703;
  p.SurfSc_gama = 1.5;
  p.SurfSc_Nbmod = 2.0;
  p.SurfSc_Npf = -0.6;
  p.SurfSc_Pft = 0.7;
  p.SurfSc_Pfn = 0.3;
  p.SurfSc_Kpha = 3600e-4;
  // src/Constant_GlobalVariable_MathFunction.cpp:19,27
  p.charge_sign[0] = -1.0; p.charge_sign[1] = 1.0; p.charge_sign[2] = -1.0;
  p.mass_eff[0] = 0.286; p.mass_eff[1] = 0.8; p.mass_eff[2] = 1.0;
}

// ---------------------------------------------------------------- helpers
void Band::out_wedge(double& xo, double& yo, double& zo, double xi, double yi, double zi,
                     int isym) const {
  // Band::OutWedge, src/Band.cpp:555-570
  const double v[3] = {xi, yi, zi};
  xo = v[matsym[0][isym] - 1] * matsym[3][isym];
  yo = v[matsym[1][isym] - 1] * matsym[4][isym];
  zo = v[matsym[2][isym] - 1] * matsym[5][isym];
}

void Band::confbz(double& x, double& y, double& z) const {
  // Band::CONFBZ, src/Band.cpp:573-616: fold into the truncated octahedron
  for (;;) {
    bool out = std::fabs(x) > a0pi || std::fabs(y) > a0pi || std::fabs(z) > a0pi ||
               std::fabs(x) + std::fabs(y) + std::fabs(z) > 1.50 * a0pi;
    if (!out) return;
    x = x > 0.0 ? x - a0pi : x + a0pi;
    y = y > 0.0 ? y - a0pi : y + a0pi;
    z = z > 0.0 ? z - a0pi : z + a0pi;
  }
}

static double fii(double x1, double y1, double z1, double x2, double y2, double z2, double x3,
                  double y3, double z3) {
  // Band::FII, src/Band.cpp:335-352: triangle area from the edge cross product
  double sbx = (y2 - y1) * (z3 - z1) - (y3 - y1) * (z2 - z1);
  double sby = (x3 - x1) * (z2 - z1) - (x2 - x1) * (z3 - z1);
  double sbz = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
  return 0.50 * std::sqrt(sbx * sbx + sby * sby + sbz * sbz);
}

double Band::surf(double e, int it) const {
  // Band::SURF, src/Band.cpp:194-332: exact iso-energy cross-section area
  const int c0 = tet_corner(0, it), c1 = tet_corner(1, it), c2 = tet_corner(2, it),
            c3 = tet_corner(3, it);
  const double e0 = eek[c0], e1 = eek[c1], e2 = eek[c2], e3 = eek[c3];
  if (!(e > e0 && e < e3)) return 0.0;
  double r;
  if (e <= e1) {
    r = (e - e0) / (e1 - e0);
    double ax = r * (xkk[c1] - xkk[c0]), ay = r * (ykk[c1] - ykk[c0]), az = r * (zkk[c1] - zkk[c0]);
    r = (e - e0) / (e2 - e0);
    double bx = r * (xkk[c2] - xkk[c0]), by = r * (ykk[c2] - ykk[c0]), bz = r * (zkk[c2] - zkk[c0]);
    r = (e - e0) / (e3 - e0);
    double cx = r * (xkk[c3] - xkk[c0]), cy = r * (ykk[c3] - ykk[c0]), cz = r * (zkk[c3] - zkk[c0]);
    return fii(ax, ay, az, bx, by, bz, cx, cy, cz);
  } else if (e > e1 && e < e2) {
    r = (e - e0) / (e2 - e0);
    double ax = r * (xkk[c2] - xkk[c0]) + xkk[c0], ay = r * (ykk[c2] - ykk[c0]) + ykk[c0],
           az = r * (zkk[c2] - zkk[c0]) + zkk[c0];
    r = (e - e1) / (e2 - e1);
    double bx = r * (xkk[c2] - xkk[c1]) + xkk[c1], by = r * (ykk[c2] - ykk[c1]) + ykk[c1],
           bz = r * (zkk[c2] - zkk[c1]) + zkk[c1];
    r = (e - e0) / (e3 - e0);
    double cx = r * (xkk[c3] - xkk[c0]) + xkk[c0], cy = r * (ykk[c3] - ykk[c0]) + ykk[c0],
           cz = r * (zkk[c3] - zkk[c0]) + zkk[c0];
    r = (e - e1) / (e3 - e1);
    double dx = r * (xkk[c3] - xkk[c1]) + xkk[c1], dy = r * (ykk[c3] - ykk[c1]) + ykk[c1],
           dz = r * (zkk[c3] - zkk[c1]) + zkk[c1];
    double a1 = fii(ax, ay, az, bx, by, bz, cx, cy, cz);
    double a2 = fii(dx, dy, dz, bx, by, bz, cx, cy, cz);
    double a3 = fii(ax, ay, az, dx, dy, dz, cx, cy, cz);
    double a4 = fii(ax, ay, az, bx, by, bz, dx, dy, dz);
    return 0.50 * (a1 + a2 + a3 + a4);
  } else if (e > e2 && e < e3) {
    r = (e - e3) / (e0 - e3);
    double ax = r * (xkk[c0] - xkk[c3]), ay = r * (ykk[c0] - ykk[c3]), az = r * (zkk[c0] - zkk[c3]);
    r = (e - e3) / (e1 - e3);
    double bx = r * (xkk[c1] - xkk[c3]), by = r * (ykk[c1] - ykk[c3]), bz = r * (zkk[c1] - zkk[c3]);
    r = (e - e3) / (e2 - e3);
    double cx = r * (xkk[c2] - xkk[c3]), cy = r * (ykk[c2] - ykk[c3]), cz = r * (zkk[c2] - zkk[c3]);
    return fii(ax, ay, az, bx, by, bz, cx, cy, cz);
  }
  return 0.0;
}

double Band::fastsurf(double e, int it) const {
  // Band::FASTSURF, src/Band.cpp:665-703: piecewise-quadratic area law
  const double e0 = eek[tet_corner(0, it)], e1 = eek[tet_corner(1, it)],
               e2 = eek[tet_corner(2, it)], e3 = eek[tet_corner(3, it)];
  double r = 0.0;
  if (e > e0 && e < e3) {
    if (e <= e1)
      r = faclow[it] * (e - e0) * (e - e0);
    else if (e < e2) {
      if (e1 - e0 > 1e-10)
        r = faclow[it] * (e - e0) * (e - e0) - facmedium[it] * (e - e1) * (e - e1);
      else
        r = surf(e, it);
    } else
      r = fachigh[it] * (e - e3) * (e - e3);
  }
  return r;
}

double Band::caldos(double e, int ib) const {
  // Band::CALDOS, src/Band.cpp:2234-2264
  const double* d = &dos[(size_t)ib * NTAB1];
  if (e <= 0e0) return d[0];
  if (e >= emax) return d[MTAB];
  int itab = (int)(e / dtable);
  double intp = (e - energy[itab]) / (energy[itab + 1] - energy[itab]);
  return d[itab] + intp * (d[itab + 1] - d[itab]);
}

double Band::caldossum(double e, int ip) const {
  // Band::CALDOSSUM, src/Band.cpp:2267-2295
  if (e <= 0e0) return sumdos[0 * 3 + ip];
  if (e >= emax) return sumdos[(size_t)MTAB * 3 + ip];
  int itab = (int)(e / dtable);
  double intp = (e - energy[itab]) / (energy[itab + 1] - energy[itab]);
  return sumdos[(size_t)itab * 3 + ip] + intp * (sumdos[(size_t)(itab + 1) * 3 + ip] - sumdos[(size_t)itab * 3 + ip]);
}

double Band::calscattsum(double e, int ib) const {
  // Band::CALSCATTSUM, src/Band.cpp:1999-2031
  if (e <= 0.0) return sumscatt[0 * MNB + ib];
  if (e >= emax) return sumscatt[(size_t)MTAB * MNB + ib];
  int itab = (int)(e / dtable);
  double intp = (e - energy[itab]) / (energy[itab + 1] - energy[itab]);
  return sumscatt[(size_t)itab * MNB + ib] +
         intp * (sumscatt[(size_t)(itab + 1) * MNB + ib] - sumscatt[(size_t)itab * MNB + ib]);
}

// ---------------------------------------------------------------- READBS & friends
static std::string read_bs(Band& b, const Phys& p, const std::string& dir) {
  // Band::READBS, src/Band.cpp:754-983 (silicon file)
  std::ifstream f(dir + "/bs.si.asc");
  if (!f) return "cannot open " + dir + "/bs.si.asc";
  for (int i = 0; i < NumParType; ++i) f >> b.nband[i];
  if (!f) return "bs.si.asc: bad header";
  if (b.nband[PELEC] > NBE || b.nband[PHOLE] > NBH || b.nband[POXEL] > NBOE)
    return "bs.si.asc: too many bands";
  b.nbt = b.nband[PELEC] + b.nband[PHOLE] + b.nband[POXEL];
  b.bandof[PELEC] = 0;
  b.bandof[PHOLE] = b.nband[PELEC];
  b.bandof[POXEL] = b.nband[PELEC] + b.nband[PHOLE];
  for (int ip = 0; ip < NumParType; ++ip)
    for (int ib = b.bandof[ip]; ib < b.bandof[ip] + b.nband[ip]; ++ib) b.partyp[ib] = ip;
  f >> b.nk;
  if (!f || b.nk <= 0 || b.nk > 1000000) return "bs.si.asc: bad nk";
  b.xkk.resize(b.nk); b.ykk.resize(b.nk); b.zkk.resize(b.nk); b.eek.resize(b.nk);
  for (auto& v : b.xkk) f >> v;
  for (auto& v : b.ykk) f >> v;
  for (auto& v : b.zkk) f >> v;
  for (auto& v : b.eek) f >> v;
  for (double e : b.eek)
    if (e < 0.0) return "bs.si.asc: negative band energy";
  f >> b.nt;
  if (!f || b.nt <= 0 || b.nt > 10000000) return "bs.si.asc: bad nt";
  const int nt = b.nt;
  b.tet.resize(4 * (size_t)nt);
  b.neib.resize(4 * (size_t)nt);
  b.ibt.resize(nt);
  int v;
  for (int c = 0; c < 4; ++c)
    for (int it = 0; it < nt; ++it) { f >> v; b.tet[(size_t)c * nt + it] = v - 1; }
  for (int c = 0; c < 4; ++c)
    for (int it = 0; it < nt; ++it) { f >> v; b.neib[(size_t)c * nt + it] = v - 1; }
  for (int it = 0; it < nt; ++it) { f >> v; b.ibt[it] = v - 1; }
  if (!f) return "bs.si.asc: truncated";
  for (int c = 0; c < 4; ++c)
    for (int it = 0; it < nt; ++it) {
      int ik = b.tet[(size_t)c * nt + it];
      if (ik < 0 || ik >= b.nk) return "bs.si.asc: corner index out of range";
    }
  for (int it = 0; it < nt; ++it)
    if (b.ibt[it] < 0 || b.ibt[it] >= b.nbt) return "bs.si.asc: band index out of range";
  for (int ik = 0; ik < b.nk; ++ik) {
    b.xkk[ik] = b.xkk[ik] * p.a0pi;
    b.ykk[ik] = b.ykk[ik] * p.a0pi;
    b.zkk[ik] = b.zkk[ik] * p.a0pi;
    b.eek[ik] = b.eek[ik] / p.eV0;
  }
  b.xkct.resize(nt); b.ykct.resize(nt); b.zkct.resize(nt);
  for (int it = 0; it < nt; ++it) {
    int c0 = b.tet_corner(0, it), c1 = b.tet_corner(1, it), c2 = b.tet_corner(2, it), c3 = b.tet_corner(3, it);
    b.xkct[it] = 0.250 * (b.xkk[c0] + b.xkk[c1] + b.xkk[c2] + b.xkk[c3]);
    b.ykct[it] = 0.250 * (b.ykk[c0] + b.ykk[c1] + b.ykk[c2] + b.ykk[c3]);
    b.zkct[it] = 0.250 * (b.zkk[c0] + b.zkk[c1] + b.zkk[c2] + b.zkk[c3]);
  }
  return "";
}

static void velotet(Band& b) {
  // Band::VELOTET, src/Band.cpp:3-62: constant group velocity solving the 3x3
  // system that reproduces the corner energies.
  const int nt = b.nt;
  b.vt.assign(4 * (size_t)nt, 0.0);
  b.vgt.assign(nt, 0.0);
  for (int it = 0; it < nt; ++it) {
    const int c0 = b.tet_corner(0, it);
    double m[4][4], mi[4][4], ar[4];
    for (int r = 1; r <= 3; ++r) {
      const int cr = b.tet_corner(r, it);
      m[r][1] = b.xkk[cr] - b.xkk[c0];
      m[r][2] = b.ykk[cr] - b.ykk[c0];
      m[r][3] = b.zkk[cr] - b.zkk[c0];
      ar[r] = b.eek[cr] - b.eek[c0];
    }
    double det = m[1][1] * (m[2][2] * m[3][3] - m[3][2] * m[2][3]) -
                 m[1][2] * (m[2][1] * m[3][3] - m[3][1] * m[2][3]) +
                 m[1][3] * (m[2][1] * m[3][2] - m[3][1] * m[2][2]);
    mi[1][1] = (m[2][2] * m[3][3] - m[3][2] * m[2][3]) / det;
    mi[2][1] = -(m[2][1] * m[3][3] - m[3][1] * m[2][3]) / det;
    mi[3][1] = (m[2][1] * m[3][2] - m[3][1] * m[2][2]) / det;
    mi[1][2] = -(m[1][2] * m[3][3] - m[3][2] * m[1][3]) / det;
    mi[2][2] = (m[1][1] * m[3][3] - m[3][1] * m[1][3]) / det;
    mi[3][2] = -(m[1][1] * m[3][2] - m[3][1] * m[1][2]) / det;
    mi[1][3] = (m[1][2] * m[2][3] - m[2][2] * m[1][3]) / det;
    mi[2][3] = -(m[1][1] * m[2][3] - m[2][1] * m[1][3]) / det;
    mi[3][3] = (m[1][1] * m[2][2] - m[2][1] * m[1][2]) / det;
    double* v = &b.vt[4 * (size_t)it];
    v[0] = mi[1][1] * ar[1] + mi[1][2] * ar[2] + mi[1][3] * ar[3];
    v[1] = mi[2][1] * ar[1] + mi[2][2] * ar[2] + mi[2][3] * ar[3];
    v[2] = mi[3][1] * ar[1] + mi[3][2] * ar[2] + mi[3][3] * ar[3];
    v[3] = b.eek[c0];
    b.vgt[it] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
}

static void hesstet(Band& b) {
  // Band::HESSTET, src/Band.cpp:356-463: inward face normals n and offsets d,
  // face f is the face opposite corner f's neighbour set as in the reference:
  // face0 = (1,2,3), face1 = (0,2,3)... (each normal built from two edges).
  const int nt = b.nt;
  b.dnt.assign(16 * (size_t)nt, 0.0);
  for (int it = 0; it < nt; ++it) {
    const int c[4] = {b.tet_corner(0, it), b.tet_corner(1, it), b.tet_corner(2, it), b.tet_corner(3, it)};
    auto X = [&](int i) { return b.xkk[c[i]]; };
    auto Y = [&](int i) { return b.ykk[c[i]]; };
    auto Z = [&](int i) { return b.zkk[c[i]]; };
    const double ax = X(1) - X(0), ay = Y(1) - Y(0), az = Z(1) - Z(0);
    const double bx = X(2) - X(1), by = Y(2) - Y(1), bz = Z(2) - Z(1);
    const double dx = X(3) - X(0), dy = Y(3) - Y(0), dz = Z(3) - Z(0);
    const double ex = X(3) - X(1), ey = Y(3) - Y(1), ez = Z(3) - Z(1);
    const double fx = X(3) - X(2), fy = Y(3) - Y(2), fz = Z(3) - Z(2);
    // {normal, point on plane, opposite corner}
    struct F { double nx, ny, nz; int on, opp; };
    const F faces[4] = {
        {ey * fz - ez * fy, ez * fx - ex * fz, ex * fy - ey * fx, 1, 0},
        {dy * fz - dz * fy, dz * fx - dx * fz, dx * fy - dy * fx, 2, 1},
        {ey * dz - ez * dy, ez * dx - ex * dz, ex * dy - ey * dx, 3, 2},
        {ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx, 0, 3},
    };
    for (int fi = 0; fi < 4; ++fi) {
      double nx = faces[fi].nx, ny = faces[fi].ny, nz = faces[fi].nz;
      const int on = faces[fi].on, opp = faces[fi].opp;
      double d = nx * X(on) + ny * Y(on) + nz * Z(on);
      if (nx * X(opp) + ny * Y(opp) + nz * Z(opp) < d) {
        nx = -nx; ny = -ny; nz = -nz; d = -d;
      }
      b.dn(0, fi, it) = nx;
      b.dn(1, fi, it) = ny;
      b.dn(2, fi, it) = nz;
      b.dn(3, fi, it) = d;
    }
  }
}

static void setmatsym(Band& b) {
  // Band::SETMATSYM, src/Band.cpp:986-1296.  The 48 operations in the reference's
  // order: rows are (axis permutation), (signs); blocks of eight share the valley
  // axis that the wedge's kx maps to (+x,-x,+y,-y,+z,-z).
  static const int perm[48][3] = {
      {1,2,3},{1,2,3},{1,2,3},{1,2,3},{1,3,2},{1,3,2},{1,3,2},{1,3,2},
      {1,2,3},{1,2,3},{1,2,3},{1,2,3},{1,3,2},{1,3,2},{1,3,2},{1,3,2},
      {2,1,3},{2,1,3},{2,1,3},{2,1,3},{3,1,2},{3,1,2},{3,1,2},{3,1,2},
      {2,1,3},{2,1,3},{2,1,3},{2,1,3},{3,1,2},{3,1,2},{3,1,2},{3,1,2},
      {2,3,1},{2,3,1},{2,3,1},{2,3,1},{3,2,1},{3,2,1},{3,2,1},{3,2,1},
      {2,3,1},{2,3,1},{2,3,1},{2,3,1},{3,2,1},{3,2,1},{3,2,1},{3,2,1}};
  static const int sg[48][3] = {
      {1,1,1},{1,1,-1},{1,-1,1},{1,-1,-1},{1,1,1},{1,1,-1},{1,-1,1},{1,-1,-1},
      {-1,1,1},{-1,1,-1},{-1,-1,1},{-1,-1,-1},{-1,1,1},{-1,1,-1},{-1,-1,1},{-1,-1,-1},
      {1,1,1},{1,1,-1},{-1,1,-1},{-1,1,1},{1,1,1},{1,1,-1},{-1,1,-1},{-1,1,1},
      {1,-1,1},{1,-1,-1},{-1,-1,-1},{-1,-1,1},{1,-1,1},{1,-1,-1},{-1,-1,-1},{-1,-1,1},
      {1,1,1},{1,-1,1},{-1,-1,1},{-1,1,1},{1,1,1},{1,-1,1},{-1,-1,1},{-1,1,1},
      {1,1,-1},{1,-1,-1},{-1,-1,-1},{-1,1,-1},{1,1,-1},{1,-1,-1},{-1,-1,-1},{-1,1,-1}};
  for (int s = 0; s < 48; ++s)
    for (int r = 0; r < 3; ++r) {
      b.matsym[r][s] = perm[s][r];
      b.matsym[3 + r][s] = sg[s][r];
    }
  for (int i = 0; i < 729; ++i) b.indmat[i] = -1;
  for (int s = 0; s < 48; ++s) {
    int idx = ((((b.matsym[3][s] + 1) * 3 + (b.matsym[4][s] + 1)) * 3 + (b.matsym[5][s] + 1)) * 3 +
               (b.matsym[0][s] - 1)) * 9 + (b.matsym[1][s] - 1) * 3 + (b.matsym[2][s] - 1);
    b.indmat[idx] = s;
  }
}

// Index of indmat[sx+1][sy+1][sz+1][p0][p1][p2].
static inline int indmat_index(int sx, int sy, int sz, int p0, int p1, int p2) {
  return ((((sx + 1) * 3 + (sy + 1)) * 3 + (sz + 1)) * 3 + p0) * 9 + p1 * 3 + p2;
}

// isym of the wedge containing (x,y,z): the ordering of |x|,|y|,|z| gives the
// permutation and the signs give the reflections (src/Band.cpp:518-545,
// src/ParticleUtils.cpp:407-430).
static int wedge_of(const Band& b, double x, double y, double z) {
  int pos[3] = {1, 1, 1}, mink, maxk;
  double tmin, tmax;
  if (std::fabs(x) > std::fabs(y)) { tmin = std::fabs(y); tmax = std::fabs(x); mink = 1; maxk = 0; }
  else { tmin = std::fabs(x); tmax = std::fabs(y); mink = 0; maxk = 1; }
  if (std::fabs(z) < tmin) mink = 2;
  else if (std::fabs(z) > tmax) maxk = 2;
  pos[mink] = 2;
  pos[maxk] = 0;
  return b.indmat[indmat_index(sgn(x), sgn(y), sgn(z), pos[0], pos[1], pos[2])];
}

static void setneibsym(Band& b) {
  // Band::SETNEIBSYM, src/Band.cpp:466-552: probe vectors just outside the five
  // faces of the irreducible wedge give the neighbouring wedge and the BZ shift.
  const double a = b.a0pi;
  const double xt[5] = {0.5 * a, 0.49 * a, 0.5 * a, 0.71 * a, 1.01 * a};
  const double yt[5] = {0.3 * a, 0.51 * a, 0.24 * a, 0.51 * a, 0.1 * a};
  const double zt[5] = {-0.01 * a, 0.1 * a, 0.26 * a, 0.31 * a, 0.05 * a};
  for (int s = 0; s < 48; ++s)
    for (int iv = 0; iv < 5; ++iv) {
      double xk, yk, zk;
      b.out_wedge(xk, yk, zk, xt[iv], yt[iv], zt[iv], s);
      const double xkl = xk, ykl = yk, zkl = zk;
      b.confbz(xk, yk, zk);
      b.xbz[iv][s] = xk - xkl;
      b.ybz[iv][s] = yk - ykl;
      b.zbz[iv][s] = zk - zkl;
      if (std::fabs(b.xbz[iv][s]) < 1e-10) b.xbz[iv][s] = 0.0;
      if (std::fabs(b.ybz[iv][s]) < 1e-10) b.ybz[iv][s] = 0.0;
      if (std::fabs(b.zbz[iv][s]) < 1e-10) b.zbz[iv][s] = 0.0;
      b.newisym[iv][s] = wedge_of(b, xk, yk, zk);
    }
}

static void calcfac(Band& b) {
  // Band::CALCFAC, src/Band.cpp:619-662
  const int nt = b.nt;
  b.faclow.assign(nt, 0.0); b.facmedium.assign(nt, 0.0); b.fachigh.assign(nt, 0.0);
  for (int it = 0; it < nt; ++it) {
    const double e0 = b.eek[b.tet_corner(0, it)], e1 = b.eek[b.tet_corner(1, it)],
                 e2 = b.eek[b.tet_corner(2, it)], e3 = b.eek[b.tet_corner(3, it)];
    double eps;
    if (e1 > e0) {
      eps = 0.50 * (e0 + e1);
      b.faclow[it] = b.surf(eps, it) / ((eps - e0) * (eps - e0));
    } else b.faclow[it] = 0.0;
    if (e3 > e2) {
      eps = 0.50 * (e2 + e3);
      b.fachigh[it] = b.surf(eps, it) / ((eps - e3) * (eps - e3));
    } else b.fachigh[it] = 0.0;
    if (e2 > e1) {
      eps = 0.5 * (e2 + e1);
      b.facmedium[it] = (b.faclow[it] * (eps - e0) * (eps - e0) - b.surf(eps, it)) / ((eps - e1) * (eps - e1));
    } else b.facmedium[it] = 0.0;
  }
}

static void getmaxaovtet(Band& b) {
  // Band::GETMAXAOVTET, src/Band.cpp:707-749: 20-step bisection on the area peak
  const int nt = b.nt;
  b.maxaovtet.assign(nt, 0.0);
  for (int it = 0; it < nt; ++it) {
    double epsl = b.eek[b.tet_corner(1, it)], epsu = b.eek[b.tet_corner(2, it)];
    double areal = b.fastsurf(epsl, it), areau = b.fastsurf(epsu, it);
    for (int is = 0; is < 20; ++is) {
      double epsm = 0.50 * (epsl + epsu);
      if (areau > areal) { epsl = epsm; areal = b.fastsurf(epsl, it); }
      else { epsu = epsm; areau = b.fastsurf(epsu, it); }
    }
    double area = areal > areau ? areal : areau;
    b.maxaovtet[it] = area / b.vgt[it];
  }
}

// Counting-sort bucket list with the reference's layout (src/Band.cpp:3843-4121):
// buckets ordered band-major then bin; pointers are filled backwards over the
// tetrahedra, so inside a bucket the tetrahedra appear in descending index.
static std::string build_bucket_list(const Band& b, int nbins, double scale, bool clamp_top,
                                     bool surface_only, std::vector<int>& n, std::vector<int>& ptr,
                                     std::vector<int>& list) {
  const int nt = b.nt;
  n.assign((size_t)nbins * MNB, 0);
  ptr.assign((size_t)nbins * MNB, 0);
  auto is_surface = [&](int it) {
    for (int d = 0; d < 4; ++d) {
      int nb = b.neib[(size_t)d * nt + it];
      if (nb == -5 || nb == -6) return true;
    }
    return false;
  };
  auto range = [&](int it, int& ib, int& ie) {
    ib = (int)(b.eek[b.tet_corner(0, it)] * scale);
    ie = (int)(b.eek[b.tet_corner(3, it)] * scale);
    if (clamp_top && ie >= nbins) ie = nbins - 1;
  };
  for (int it = 0; it < nt; ++it) {
    int ib, ie;
    range(it, ib, ie);
    if (!clamp_top && ie >= nbins) return "tetrahedron energy beyond the coarse list range";
    if (ib < 0) continue;
    if (surface_only && !is_surface(it)) continue;
    for (int t = ib; t <= ie; ++t) n[(size_t)t * MNB + b.ibt[it]] += 1;
  }
  int p = 0;
  for (int ib = 0; ib < b.nbt; ++ib)
    for (int t = 0; t < nbins; ++t) {
      p += n[(size_t)t * MNB + ib];
      ptr[(size_t)t * MNB + ib] = p;
    }
  list.assign(p > 0 ? p : 1, -1);
  for (int it = 0; it < nt; ++it) {
    int ib, ie;
    range(it, ib, ie);
    if (ib < 0) continue;
    if (surface_only && !is_surface(it)) continue;
    for (int t = ib; t <= ie; ++t) {
      int& q = ptr[(size_t)t * MNB + b.ibt[it]];
      q -= 1;
      list[q] = it;
    }
  }
  return "";
}

static std::string buildlists(Band& b, const Phys& p) {
  // Band::BUILDLISTS, src/Band.cpp:3843-4121.  The cube list (clist) is left out:
  // OVERLAP's entry guard (src/Band.cpp:2355-2372) only admits tetrahedra whose
  // corners all have ky >= 0.075 a0pi and kx <= 0.7 a0pi, for which the cube
  // ranges icy >= 100 > icymax, so every cube list is empty and the cube search
  // in Getisym (src/ParticleUtils.cpp:450-480) never finds a tetrahedron.
  // The Fischetti lists are only read when fiscphfl is set, which IELEC clears.
  b.dlist = p.eV0 / 0.040;
  b.dlists = p.eV0 / 0.0020;
  std::string err = build_bucket_list(b, MWLE, b.dlist, false, false, b.ntlist, b.ptlist, b.tlist);
  if (!err.empty()) return err;
  err = build_bucket_list(b, MWLE, b.dlist, false, true, b.nslist, b.pslist, b.slist);
  if (!err.empty()) return err;
  return build_bucket_list(b, MWLES, b.dlists, true, false, b.ntlists, b.ptlists, b.tlists);
}

static std::string getdos(Band& b, const Phys& p, const std::string& dir) {
  // Band::GETDOS, src/Band.cpp:3661-3834 (tables read from zd.si.asc)
  std::ifstream f(dir + "/zd.si.asc");
  if (!f) return "cannot open " + dir + "/zd.si.asc";
  int itab;
  double dum;
  f >> itab;
  if (!f || itab != MTAB) return "zd.si.asc: wrong number of table entries";
  b.dos.assign((size_t)MNB * NTAB1, 0.0);
  b.sumdos.assign((size_t)NTAB1 * 3, 0.0);
  b.dostetmax.assign((size_t)MNB * NTAB1, 0.0);
  for (int t = 0; t <= MTAB; ++t) {
    f >> dum;
    for (int ib = 0; ib < b.nbt; ++ib) f >> b.dos[(size_t)ib * NTAB1 + t];
    for (int ip = 0; ip < NumParType; ++ip) {
      double s = 0.0;
      for (int ib = b.bandof[ip]; ib < b.nband[ip] + b.bandof[ip]; ++ib) s = s + b.dos[(size_t)ib * NTAB1 + t];
      b.sumdos[(size_t)t * 3 + ip] = s;
    }
  }
  f >> itab;
  for (int t = 0; t <= MTAB; ++t) {
    f >> dum;
    for (int ib = 0; ib < b.nbt; ++ib) f >> b.dostetmax[(size_t)ib * NTAB1 + t];
  }
  if (!f) return "zd.si.asc: truncated";
  for (int ip = 0; ip < NumParType; ++ip) b.DOSMAX[ip] = 0.0;
  const double s3 = std::pow(p.spr0, 3.0);
  for (int t = 0; t <= MTAB; ++t) {
    for (int ib = 0; ib < b.nbt; ++ib) {
      b.dos[(size_t)ib * NTAB1 + t] = b.dos[(size_t)ib * NTAB1 + t] * p.eV0 * s3;
      b.dostetmax[(size_t)ib * NTAB1 + t] = b.dostetmax[(size_t)ib * NTAB1 + t] * p.velo0 / (p.spk0 * p.spk0);
    }
    for (int ip = 0; ip < NumParType; ++ip) {
      b.sumdos[(size_t)t * 3 + ip] = b.sumdos[(size_t)t * 3 + ip] * p.eV0 * s3;
      if (b.sumdos[(size_t)t * 3 + ip] > b.DOSMAX[ip]) b.DOSMAX[ip] = b.sumdos[(size_t)t * 3 + ip];
    }
  }
  return "";
}

static double eiirate(const Band& b, const Phys& p, double ee) {
  // Band::EIIRATE, src/Band.cpp:1913-1971 with only the Thoma model enabled
  double eel = ee - p.sieg + 1.1241e0 / p.eV0;
  double sii = 0.0;
  if ((eel * p.eV0 > 1.128e0) && (eel * p.eV0 < 1.75e0))
    sii = 1.25e12 / p.scrt0 * std::pow((eel * p.eV0 - 1.128e0), 3);
  else if (eel * p.eV0 > 1.75e0)
    sii = 9.49e12 / p.scrt0 * (eel * p.eV0 - 1.572e0) * (eel * p.eV0 - 1.572e0);
  return sii * b.iifacelec;
}

static double hiirate(const Band& b, const Phys& p, double ee) {
  // Band::HIIRATE, src/Band.cpp:1975-1996
  double eel = ee - p.sieg + 1.1241e0 / p.eV0;
  if (eel > b.hiithresh) return b.iifachole * std::pow((eel - b.hiithresh), b.hiiexp);
  return 0.0;
}

static void buildphscatt(Band& b, const Phys& p) {
  // Band::BUILDPHSCATT, src/Band.cpp:1302-1910, Jacoboni phonon system
  const double temps[6] = {b.temptag, b.templag, b.templog, b.temptaf, b.templaf, b.temptof};
  double bes[6];
  for (int i = 0; i < 6; ++i) bes[i] = 1.0 / (std::exp(temps[i]) - 1.0);
  b.scpre = 14;
  const double kint = 2.0 * PI * b.dfelast * b.dfelast / p.sirho / 6e0 / (p.siul * p.siul);
  double kem[6], kab[6];
  kem[0] = (b.dftag) * (b.dftag) * PI / b.temptag / p.sirho / 6e0;
  kem[1] = (b.dflag) * (b.dflag) * PI / b.templag / p.sirho / 6e0;
  kem[2] = (b.dflog) * (b.dflog) * PI / b.templog / p.sirho / 6e0;
  kem[3] = (b.dftaf) * (b.dftaf) * PI / b.temptaf / p.sirho * 0.666e0;
  kem[4] = (b.dflaf) * (b.dflaf) * PI / b.templaf / p.sirho * 0.666e0;
  kem[5] = (b.dftof) * (b.dftof) * PI / b.temptof / p.sirho * 0.666e0;
  for (int i = 0; i < 6; ++i) kab[i] = bes[i] * kem[i];
  for (int i = 0; i < 6; ++i) kem[i] = (bes[i] + 1.0) * kem[i];

  b.scatte.assign((size_t)MNScaEle * NBE * NBE, 0.0);
  b.dose.assign((size_t)MNScaEle * NBE * NTAB1, 0.0);
  auto DOSE = [&](int s, int f, int t) -> double& { return b.dose[((size_t)s * NBE + f) * NTAB1 + t]; };
  auto SCE = [&](int s, int f, int i) -> double& { return b.scatte[((size_t)s * NBE + f) * NBE + i]; };
  int ipt = PELEC;
  for (int fin = 0; fin < b.nband[ipt]; ++fin)
    for (int t = 0; t <= MTAB; ++t) {
      const double ee = b.energy[t];
      const int ib = fin + b.bandof[ipt];
      DOSE(0, fin, t) = b.caldos(ee, ib);
      for (int m = 0; m < 6; ++m) {
        DOSE(1 + 2 * m, fin, t) = b.caldos(ee + temps[m], ib);  // absorption
        DOSE(2 + 2 * m, fin, t) = b.caldos(ee - temps[m], ib);  // emission
      }
    }
  for (int ini = 0; ini < b.nband[ipt]; ++ini)
    for (int fin = 0; fin < b.nband[ipt]; ++fin) {
      const double dosfac = (ini == 0 && fin == 0) ? 1.0 : b.ephb;
      SCE(0, fin, ini) = kint * dosfac;
      for (int m = 0; m < 6; ++m) {
        SCE(1 + 2 * m, fin, ini) = kab[m] * dosfac;
        SCE(2 + 2 * m, fin, ini) = kem[m] * dosfac;
      }
    }
  // holes
  b.scprh = 4;
  const double besh = 1.0 / (std::exp(b.temphop) - 1.0);
  const double kinth = 2e0 * PI * b.dfhelast * b.dfhelast / p.sirho / 2.0 / (p.siul * p.siul);
  double kemh = b.dfhop * b.dfhop * PI / b.temphop / p.sirho;
  const double kabh = besh * kemh;
  kemh = (besh + 1.0) * kemh;
  b.scatth.assign((size_t)MNScaHole * NBH * NBH, 0.0);
  b.dosh.assign((size_t)MNScaHole * NBH * NTAB1, 0.0);
  auto DOSH = [&](int s, int f, int t) -> double& { return b.dosh[((size_t)s * NBH + f) * NTAB1 + t]; };
  auto SCH = [&](int s, int f, int i) -> double& { return b.scatth[((size_t)s * NBH + f) * NBH + i]; };
  ipt = PHOLE;
  for (int fin = 0; fin < b.nband[ipt]; ++fin)
    for (int t = 0; t <= MTAB; ++t) {
      const double ee = b.energy[t];
      const int ib = fin + b.bandof[ipt];
      DOSH(0, fin, t) = b.caldos(ee, ib);
      DOSH(1, fin, t) = b.caldos(ee + b.temphop, ib);
      DOSH(2, fin, t) = b.caldos(ee - b.temphop, ib);
    }
  for (int ini = 0; ini < b.nband[ipt]; ++ini)
    for (int fin = 0; fin < b.nband[ipt]; ++fin) {
      SCH(0, fin, ini) = kinth;
      SCH(1, fin, ini) = kabh;
      SCH(2, fin, ini) = kemh;
    }
  // impact ionisation (Taniguchi final energy (E - Eg)/3)
  b.scattiie.assign((size_t)NBE * NBE * NTAB1, 0.0);
  b.scattiih.assign((size_t)NBH * NBH * NTAB1, 0.0);
  for (int ptype = PELEC; ptype <= PHOLE; ++ptype) {
    const int nb = b.nband[ptype];
    const int nbmax = ptype == PELEC ? NBE : NBH;
    std::vector<double>& tab = ptype == PELEC ? b.scattiie : b.scattiih;
    for (int ini = 0; ini < nb; ++ini)
      for (int fin = 0; fin < nb; ++fin)
        for (int t = 0; t <= MTAB; ++t) {
          const double ee = b.energy[t];
          const double sii = ptype == PELEC ? eiirate(b, p, ee) : hiirate(b, p, ee);
          const double eeafter = (ee - p.sieg) / 3.0;
          double d1, d2;
          if (eeafter > 0.0) {
            d1 = b.caldos(eeafter, fin + b.bandof[ptype]);
            d2 = b.caldossum(eeafter, ptype);
            if (d2 == 0.0) { d1 = 0.0; d2 = 1.0; }
          } else { d1 = 0.0; d2 = 1.0; }
          tab[((size_t)fin * nbmax + ini) * NTAB1 + t] = d1 / d2 * sii;
        }
  }
  // oxide electrons (only their band's rate sum is consumed)
  b.scproe = 3;
  const double besoe = 1.0 / (std::exp(b.tempoeop) - 1.0);
  const double kintoe = 2.0 * PI * (b.dfoeelast * b.dfoeelast) / p.sirho / 2e0 / (p.siul * p.siul);
  double kemoe = (b.dfoeop * b.dfoeop) * 1.5e0 * PI / b.tempoeop / p.sirho;
  const double kaboe = besoe * kemoe;
  kemoe = (besoe + 1.0) * kemoe;
  (void)kemoe;
  // sum of rates per initial band (src/Band.cpp:1810-1888)
  b.sumscatt.assign((size_t)NTAB1 * MNB, 0.0);
  for (int t = 0; t <= MTAB; ++t) {
    ipt = PELEC;
    for (int ini = 0; ini < b.nband[ipt]; ++ini) {
      double& s = b.sumscatt[(size_t)t * MNB + ini + b.bandof[ipt]];
      for (int fin = 0; fin < b.nband[ipt]; ++fin) {
        for (int is = 0; is < b.scpre - 1; ++is) s = s + SCE(is, fin, ini) * DOSE(is, fin, t);
        s = s + b.scattiie[((size_t)fin * NBE + ini) * NTAB1 + t];
      }
    }
  }
  for (int t = 0; t <= MTAB; ++t) {
    ipt = PHOLE;
    for (int ini = 0; ini < b.nband[ipt]; ++ini) {
      double& s = b.sumscatt[(size_t)t * MNB + ini + b.bandof[ipt]];
      for (int fin = 0; fin < b.nband[ipt]; ++fin) {
        for (int is = 0; is < b.scprh - 1; ++is) s = s + SCH(is, fin, ini) * DOSH(is, fin, t);
        s = s + b.scattiih[((size_t)fin * NBH + ini) * NTAB1 + t];
      }
    }
  }
  for (int t = 0; t <= MTAB; ++t) {
    ipt = POXEL;
    const double ee = b.energy[t];
    for (int ini = 0; ini < b.nband[ipt]; ++ini) {
      double& s = b.sumscatt[(size_t)t * MNB + ini + b.bandof[ipt]];
      for (int fin = 0; fin < b.nband[ipt]; ++fin) {
        const int ib = fin + b.bandof[ipt];
        s = s + kintoe * b.caldos(ee, ib);
        s = s + kaboe * b.caldos(ee + b.tempoeop, ib);
      }
    }
  }
  // variable-Gamma bound per tetrahedron (src/Band.cpp:1890-1903)
  b.gamtet.assign(b.nt, 0.0);
  for (int it = 0; it < b.nt; ++it) {
    const int ib = b.ibt[it];
    int itl = (int)(b.eek[b.tet_corner(0, it)] / b.dtable);
    if (itl < 0) itl = 0;
    int itu = (int)(b.eek[b.tet_corner(3, it)] / b.dtable) + 1;
    if (itu > MTAB) itu = MTAB;
    double g = 0.0;
    for (int t = itl; t <= itu; ++t)
      if (b.sumscatt[(size_t)t * MNB + ib] > g) g = b.sumscatt[(size_t)t * MNB + ib];
    if (g == 0.0) g = b.sumscatt[(size_t)MTAB * MNB + ib];
    b.gamtet[it] = g;
  }
}

static void getindens(Band& b, const Phys& p) {
  // Band::GETINDENS, src/Band.cpp:3609-3656: Boltzmann integral of the DOS
  double en0 = 0.0, ep0 = 0.0;
  double eelmo = b.energy[0];
  for (int t = 1; t <= MTAB; ++t) {
    const double eel = b.energy[t];
    double dosl = b.caldossum(eel, PELEC), doslmo = b.caldossum(eelmo, PELEC);
    en0 = en0 + ((doslmo * eel - dosl * eelmo) * (std::exp(-eelmo) - std::exp(-eel)) +
                 (dosl - doslmo) * ((eelmo + 1e0) * std::exp(-eelmo) - (eel + 1e0) * std::exp(-eel))) /
                    (eel - eelmo);
    dosl = b.caldossum(eel, PHOLE);
    doslmo = b.caldossum(eelmo, PHOLE);
    ep0 = ep0 + ((doslmo * eel - dosl * eelmo) * (std::exp(-eelmo) - std::exp(-eel)) +
                 (dosl - doslmo) * ((eelmo + 1e0) * std::exp(-eelmo) - (eel + 1e0) * std::exp(-eel))) /
                    (eel - eelmo);
    eelmo = eel;
  }
  en0 = 2e0 * en0;
  ep0 = 2e0 * ep0;
  b.Ni = std::sqrt(en0 * ep0 * std::exp(-p.sieg));
}

std::string band_build(Band& b, const Phys& p, const std::string& dir) {
  // Band::IELEC, src/Band.cpp:4271-4476
  b.a0pi = p.a0pi;
  b.sieg = p.sieg;
  b.dtable = 0.0010 / p.eV0;
  b.emin = 0.0;
  b.emax = (double)(MTAB)*b.dtable;
  b.energy.resize(NTAB1);
  for (int t = 0; t <= MTAB; ++t) b.energy[t] = (double)(t)*b.dtable + b.emin;
  const double T0 = p.T0;
  b.temptag = 140.0 / T0; b.templag = 215.0 / T0; b.templog = 720.0 / T0;
  b.temptaf = 220.0 / T0; b.templaf = 550.0 / T0; b.temptof = 685.0 / T0;
  b.dftag = 4.65e9 / p.dpc0; b.dflag = 7.44e9 / p.dpc0; b.dflog = 1.023e11 / p.dpc0;
  b.dftaf = 2.79e9 / p.dpc0; b.dflaf = 1.86e10 / p.dpc0; b.dftof = 1.86e10 / p.dpc0;
  b.dfelast = 8.7 / p.eV0;
  b.temphop = 680.0 / T0; b.dfhop = 4e10 / p.dpc0; b.dfhelast = 9.3 / p.eV0;
  b.tempoeop = 735.0 / T0; b.dfoeop = 2.2e10 / p.dpc0; b.dfoeelast = 25.0 / p.eV0;
  b.iifacelec = 0.16;
  b.difpr[PELEC] = 0.16; b.difpr[PHOLE] = 0.35; b.difpr[POXEL] = 0.16;
  b.iifachole = 1.14e12;
  b.hiithresh = 1.49 / p.eV0;
  b.hiiexp = 3.4;
  b.iifachole = b.iifachole * std::pow(p.eV0, b.hiiexp) * p.time0;
  b.ephb = 1.6;
  b.mell = 0.9116;
  b.melt = 0.1946;
  b.meld = std::pow((b.mell * b.melt * b.melt), (1.0 / 3.0));

  std::string err = read_bs(b, p, dir);
  if (!err.empty()) return err;
  velotet(b);
  hesstet(b);
  setmatsym(b);
  setneibsym(b);
  calcfac(b);
  getmaxaovtet(b);
  err = buildlists(b, p);
  if (!err.empty()) return err;
  err = getdos(b, p, dir);
  if (!err.empty()) return err;
  buildphscatt(b, p);
  getindens(b, p);
  for (int ip = 0; ip < NumParType; ++ip) {
    b.gamma[ip] = 0.0;
    if (b.nband[ip] == 0) continue;
    for (int t = 0; t <= MTAB; ++t) {
      double s = b.calscattsum(b.energy[t], b.bandof[ip]);
      if (s > b.gamma[ip]) b.gamma[ip] = s;
    }
  }
  return "";
}

std::string load_phonon_tables(Band& b, const std::string& dir, const std::string files[PTAB_N]) {
  for (int t = 0; t < PTAB_N; ++t) {
    const std::string path = dir + "/" + files[t];
    std::ifstream f(path);
    if (!f) return "cannot open dispersion table " + path;
    const int R = PTAB_ROWS[t], C = PTAB_COLS[t];
    b.ptab[t].assign((size_t)R * C, 0.0);
    b.ptab_t[t].assign((size_t)R * C, 0.0);
    for (int c = 0; c < C; ++c)
      for (int r = 0; r < R; ++r) {
        double v;
        if (!(f >> v)) return "dispersion table " + path + " is short";
        b.ptab[t][(size_t)r * C + c] = v;
        b.ptab_t[t][(size_t)c * R + r] = v;
      }
  }
  return "";
}

}  // namespace emc
