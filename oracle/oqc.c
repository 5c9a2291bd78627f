/*
 * oqc.c - CPU restatement of the quantum-correction effective potential,
 * test infrastructure only:
 *   init_p_mat              src/mcmodel.cpp:2495-2582 (node material and edge flags)
 *   init_effective_potential src/mcmodel.cpp:2584-2625 (per-axis node windows)
 *   init_ep_range           src/mcmodel.cpp:2406-2493 (control-box edges, Eb)
 *   effective_potential     src/mcmodel.cpp:2337-2376
 *   quantumPotential        src/mcmodel.cpp:2378-2405
 * Per node and term, as the reference evaluates it: erf of the box edges
 * relative to the target, multiplied z, y, x after the potential, summed over
 * the window in i, j, k order, times 1/8.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define N_SILICON 1
#define N_OXIDE 2
#define N_UP (1 << 2)
#define N_DOWN (1 << 3)
#define N_FRONT (1 << 4)
#define N_BACK (1 << 5)
#define N_LEFT (1 << 6)
#define N_RIGHT (1 << 7)
#define N_QUANTUM (1 << 8)
#define MAT_OXIDE 2
#define MAT_SILICON 3

static const double SQRT2_ = 1.4142135623730950488e0; /* include/Constant_GlobalVariable_MathFunction.h:21 */

/* The reference walks two iterators over the grid lines: lower_bound over
 * [previous, p) for l[p] - r, upper_bound over [previous, end) for l[p] + r;
 * the lower one steps back one line (when it can), the upper one is clamped to
 * the last line, and both searches resume there. */
static void window(const double *l, int n, double r, int *lo, int *hi) {
  int a = 0, b = 0;
  for (int p = 0; p < n; ++p) {
    const double lv = l[p] - r, uv = l[p] + r;
    while (a < p && l[a] < lv) ++a;
    while (b < n && !(l[b] > uv)) ++b;
    if (a != 0) --a;
    if (b == n) --b;
    lo[p] = a;
    hi[p] = b;
  }
}

int orc_qc_potential(const emc_mesh_view *m, const double *pot, double *out) {
  const int cnx = m->cnx, cny = m->cny, cnz = m->cnz, pnx = m->pnx, pny = m->pny, pnz = m->pnz;
  const size_t nn = (size_t)m->nnode;
#define ND(i, j, k) (((size_t)(j) * pnx + (i)) * pnz + (k))
#define CL(i, j, k) (((size_t)(j) * cnx + (i)) * cnz + (k))
  int *pmat = (int *)calloc(nn, sizeof(int));
  int *mat = (int *)malloc(nn * sizeof(int));
  double *rg[6], *eb = (double *)calloc(nn, sizeof(double));
  for (int q = 0; q < 6; ++q) rg[q] = (double *)malloc(nn * sizeof(double));
  int *lo[3], *hi[3];
  double *ql[3];
  const int np[3] = {pnx, pny, pnz};
  const double *L[3] = {m->lx, m->ly, m->lz};
  for (int a = 0; a < 3; ++a) {
    lo[a] = (int *)malloc(np[a] * sizeof(int));
    hi[a] = (int *)malloc(np[a] * sizeof(int));
    ql[a] = (double *)malloc(np[a] * sizeof(double));
  }
  if (!pmat || !mat || !eb) return -1;
  /* init_p_mat: each cell marks its 8 corner nodes */
  for (int i = 0; i < cnx; ++i)
    for (int k = 0; k < cnz; ++k)
      for (int j = 0; j < cny; ++j) {
        int f = 0;
        if (m->qcregion[CL(i, j, k)]) f |= N_QUANTUM;
        if (m->material[CL(i, j, k)] == MAT_SILICON) f |= N_SILICON;
        if (m->material[CL(i, j, k)] == MAT_OXIDE) f |= N_OXIDE;
        for (int d = 0; d < 8; ++d) pmat[ND(i + (d & 1), j + ((d >> 1) & 1), k + (d >> 2))] |= f;
      }
  memcpy(mat, pmat, nn * sizeof(int));
#define SOLID(i, j, k) ((mat[ND(i, j, k)] & (N_OXIDE | N_SILICON)) != 0)
  for (int i = 0; i < pnx; ++i)
    for (int j = 0; j < pny; ++j)
      for (int k = 0; k < pnz; ++k) {
        int *f = &pmat[ND(i, j, k)];
        if (i == 0 || !SOLID(i - 1, j, k)) *f |= N_UP;
        if (i == pnx - 1 || !SOLID(i + 1, j, k)) *f |= N_DOWN;
        if (k == 0 || !SOLID(i, j, k - 1)) *f |= N_FRONT;
        if (k == pnz - 1 || !SOLID(i, j, k + 1)) *f |= N_BACK;
        if (j == 0 || !SOLID(i, j - 1, k)) *f |= N_LEFT;
        if (j == pny - 1 || !SOLID(i, j + 1, k)) *f |= N_RIGHT;
      }
  /* init_ep_range */
  const double xr = m->qc_ratio[0], xt = m->qc_theta[0], yr = m->qc_ratio[1], yt = m->qc_theta[1];
  const double zr = m->qc_ratio[2], zt = m->qc_theta[2];
  for (int j = 0; j < pny; ++j)
    for (int i = 0; i < pnx; ++i)
      for (int k = 0; k < pnz; ++k) {
        const size_t n = ND(i, j, k);
        const int f = pmat[n];
        if ((f & N_OXIDE) && !(f & N_SILICON)) eb[n] = -m->qc_eb;
        const double x11 = (f & N_UP) ? m->lx[i] - xr * xt : m->lx[i] - 0.5 * m->dx[i - 1];
        const double x22 = (f & N_DOWN) ? m->lx[i] + xr * xt : m->lx[i] + 0.5 * m->dx[i];
        const double y11 = (f & N_LEFT) ? m->ly[j] - yr * yt : m->ly[j] - 0.5 * m->dy[j - 1];
        const double y22 = (f & N_RIGHT) ? m->ly[j] + yr * yt : m->ly[j] + 0.5 * m->dy[j];
        const double z11 = (f & N_FRONT) ? m->lz[k] - zr * zt : m->lz[k] - 0.5 * m->dz[k - 1];
        const double z22 = (f & N_BACK) ? m->lz[k] + zr * zt : m->lz[k] + 0.5 * m->dz[k];
        rg[0][n] = x11 / (SQRT2_ * xt);
        rg[1][n] = x22 / (SQRT2_ * xt);
        rg[2][n] = y11 / (SQRT2_ * yt);
        rg[3][n] = y22 / (SQRT2_ * yt);
        rg[4][n] = z11 / (SQRT2_ * zt);
        rg[5][n] = z22 / (SQRT2_ * zt);
      }
  /* init_effective_potential */
  const double th[3] = {xt, yt, zt}, ra[3] = {xr, yr, zr};
  for (int a = 0; a < 3; ++a) {
    window(L[a], np[a], ra[a] * th[a], lo[a], hi[a]);
    for (int p = 0; p < np[a]; ++p) ql[a][p] = L[a][p] / (SQRT2_ * th[a]);
  }
  /* quantumPotential: sources are the Poisson potential as given */
  memcpy(out, pot, nn * sizeof(double));
  for (int pj = 0; pj < pny; ++pj)
    for (int pi = 0; pi < pnx; ++pi)
      for (int pk = 0; pk < pnz; ++pk) {
        if (!(pmat[ND(pi, pj, pk)] & N_QUANTUM)) continue;
        double ret = 0;
        for (int i = lo[0][pi]; i <= hi[0][pi]; ++i)
          for (int j = lo[1][pj]; j <= hi[1][pj]; ++j)
            for (int k = lo[2][pk]; k <= hi[2][pk]; ++k) {
              const size_t s = ND(i, j, k);
              if (!((pmat[s] & N_OXIDE) || (pmat[s] & N_SILICON))) continue;
              const double tmp = pot[s] + eb[s];
              const double x1 = rg[0][s] - ql[0][pi], x2 = rg[1][s] - ql[0][pi];
              const double y1 = rg[2][s] - ql[1][pj], y2 = rg[3][s] - ql[1][pj];
              const double z1 = rg[4][s] - ql[2][pk], z2 = rg[5][s] - ql[2][pk];
              ret += tmp * (erf(z2) - erf(z1)) * (erf(y2) - erf(y1)) * (erf(x2) - erf(x1));
            }
        ret *= 0.125;
        out[ND(pi, pj, pk)] = ret;
      }
  free(pmat);
  free(mat);
  free(eb);
  for (int q = 0; q < 6; ++q) free(rg[q]);
  for (int a = 0; a < 3; ++a) {
    free(lo[a]);
    free(hi[a]);
    free(ql[a]);
  }
  return 0;
#undef ND
#undef CL
#undef SOLID
}
