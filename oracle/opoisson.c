/*
 * opoisson.c - FP64 Jacobi-preconditioned CG on the Poisson operator of
 * init_matrix (src/mcmodel.cpp:4695-4869) after symmetric elimination of the
 * contact rows.  It stands in for the Amesos-KLU solve (src/PoissonSolver.cpp:70-78),
 * which is not available here; agreement is measured by the residual of the
 * operator as assembled (orc_apply_full).  Test infrastructure only.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void apply_elim(const emc_mesh_view *m, const double *x, double *y) {
  const int pnx = m->pnx, pny = m->pny, pnz = m->pnz;
  const size_t sx = (size_t)pnz, sy = (size_t)pnx * pnz;
  for (int j = 0; j < pny; ++j)
    for (int i = 0; i < pnx; ++i)
      for (int k = 0; k < pnz; ++k) {
        const size_t n = (size_t)j * sy + (size_t)i * sx + k;
        double v = m->diag[n] * x[n];
        if (i + 1 < pnx) v += m->cx[n] * x[n + sx];
        if (i > 0) v += m->cx[n - sx] * x[n - sx];
        if (j + 1 < pny) v += m->cy[n] * x[n + sy];
        if (j > 0) v += m->cy[n - sy] * x[n - sy];
        if (k + 1 < pnz) v += m->cz[n] * x[n + 1];
        if (k > 0) v += m->cz[n - 1] * x[n - 1];
        y[n] = v;
      }
}

void orc_apply_full(const emc_mesh_view *m, const double *x, double *y) {
  const int pnx = m->pnx, pny = m->pny, pnz = m->pnz;
  const size_t sx = (size_t)pnz, sy = (size_t)pnx * pnz;
  for (int j = 0; j < pny; ++j)
    for (int i = 0; i < pnx; ++i)
      for (int k = 0; k < pnz; ++k) {
        const size_t n = (size_t)j * sy + (size_t)i * sx + k;
        const double *s = &m->stencil[n * 7];
        double v = s[6] * x[n];
        if (i > 0) v += s[0] * x[n - sx];
        if (i + 1 < pnx) v += s[1] * x[n + sx];
        if (j > 0) v += s[2] * x[n - sy];
        if (j + 1 < pny) v += s[3] * x[n + sy];
        if (k > 0) v += s[4] * x[n - 1];
        if (k + 1 < pnz) v += s[5] * x[n + 1];
        y[n] = v;
      }
}

int orc_pcg(const emc_mesh_view *m, const double *b, double *x, double rtol, int maxit, double *relres) {
  const size_t N = (size_t)m->nnode;
  double *r = malloc(N * 8), *z = malloc(N * 8), *p = malloc(N * 8), *q = malloc(N * 8);
  for (size_t n = 0; n < N; ++n)
    if (m->node_kind[n] != 0) x[n] = b[n];
  apply_elim(m, x, q);
  double bn = 0, rz = 0, rn;
  for (size_t n = 0; n < N; ++n) {
    r[n] = b[n] - q[n];
    z[n] = r[n] / m->diag[n];
    p[n] = z[n];
    rz += r[n] * z[n];
    bn += b[n] * b[n];
  }
  bn = sqrt(bn);
  if (bn == 0) bn = 1;
  int it = 0;
  rn = 0;
  for (size_t n = 0; n < N; ++n) rn += r[n] * r[n];
  rn = sqrt(rn);
  while (rn / bn > rtol && it < maxit) {
    apply_elim(m, p, q);
    double pq = 0;
    for (size_t n = 0; n < N; ++n) pq += p[n] * q[n];
    const double alpha = rz / pq;
    double rr = 0, rz1 = 0;
    for (size_t n = 0; n < N; ++n) {
      x[n] += alpha * p[n];
      r[n] -= alpha * q[n];
      z[n] = r[n] / m->diag[n];
      rr += r[n] * r[n];
      rz1 += r[n] * z[n];
    }
    const double beta = rz1 / rz;
    rz = rz1;
    for (size_t n = 0; n < N; ++n) p[n] = z[n] + beta * p[n];
    rn = sqrt(rr);
    ++it;
  }
  if (relres) *relres = rn / bn;
  free(r); free(z); free(p); free(q);
  return it;
}
