/*
 * odeposit.c - restatement of statistic (src/mcmodel.cpp:1167-1408),
 * particle_to_density (src/mcmodel.cpp:348-417), compute_cell_charge
 * (src/mcmodel.cpp:2664-2712) and the RHS / field / surface-rate steps of
 * compute_field (src/mcmodel.cpp:2239-2263, 2732-2893,
 * src/GetSurfRoughnessPhononScRate.cpp:4-71).  Test infrastructure only.
 *
 * Sums are int64 fixed point: every contribution is rounded to the nearest
 * multiple of 1/scale (EMC_FX_*) before it is added, so the totals do not
 * depend on the order particles are visited.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define PI 3.141592653589793238462643e0
#define BOLTZ 8.617343e-5

static int64_t fx(double v, double scale) { return (int64_t)rint(v * scale); }

int orc_deposit(const emc_band_view *b, const emc_mesh_view *m, const emc_particles *p, int64_t n,
                int64_t *node_q, int64_t *node_stat, int64_t *cell_q, int64_t *cell_n) {
  const size_t NN = (size_t)m->nnode, NC = (size_t)m->ncell;
  if (node_q) memset(node_q, 0, NN * 8);
  if (node_stat) memset(node_stat, 0, 10 * NN * 8);
  if (cell_q) memset(cell_q, 0, 3 * NC * 8);
  if (cell_n) memset(cell_n, 0, NC * 8);
#define NODE(i, j, k) (((size_t)(j) * m->pnx + (i)) * m->pnz + (k))
  for (int64_t q = 0; q < n; ++q) {
    const int i = p->i[q], j = p->j[q], k = p->k[q];
    if (i < 0) continue;
    const size_t c = ((size_t)j * m->cnx + i) * m->cnz + k;
    const double x = p->x[q], y = p->y[q], z = p->z[q], charge = p->charge[q], energy = p->energy[q];
    const int itet = p->itet[q], isym = p->isym[q], pt = p->ptype[q];
    const double rx = (x - m->lx[i]) / m->dx[i];
    const double ry = (y - m->ly[j]) / m->dy[j];
    const double rz = (z - m->lz[k]) / m->dz[k];
    if (cell_n) cell_n[c] += 1;
    /* statistic: |q|-weighted velocity / charge / energy per node */
    if (node_stat && (pt == 0 || pt == 1)) {
      const double *vt = &b->vt[4 * (size_t)itet];
      const double v3[3] = {vt[0], vt[1], vt[2]};
      const double vx = v3[b->matsym[0 * 48 + isym] - 1] * b->matsym[3 * 48 + isym];
      const double vy = v3[b->matsym[1 * 48 + isym] - 1] * b->matsym[4 * 48 + isym];
      const double vz = v3[b->matsym[2 * 48 + isym] - 1] * b->matsym[5 * 48 + isym];
      int64_t *S = node_stat + (size_t)(pt == 0 ? 0 : 5) * NN;
      const size_t nodes[8] = {NODE(i, j, k), NODE(i + 1, j, k), NODE(i, j, k + 1), NODE(i + 1, j, k + 1),
                               NODE(i, j + 1, k), NODE(i + 1, j + 1, k), NODE(i, j + 1, k + 1),
                               NODE(i + 1, j + 1, k + 1)};
      const double cc[8] = {charge * (1 - rx) * (1 - ry) * (1 - rz), charge * rx * (1 - ry) * (1 - rz),
                            charge * (1 - rx) * (1 - ry) * rz,       charge * rx * (1 - ry) * rz,
                            charge * (1 - rx) * ry * (1 - rz),       charge * rx * ry * (1 - rz),
                            charge * (1 - rx) * ry * rz,             charge * rx * ry * rz};
      for (int a = 0; a < 8; ++a) {
        const double f = fabs(cc[a]);
        S[0 * NN + nodes[a]] += fx(vx * f, EMC_FX_STAT);
        S[1 * NN + nodes[a]] += fx(vy * f, EMC_FX_STAT);
        S[2 * NN + nodes[a]] += fx(vz * f, EMC_FX_STAT);
        S[3 * NN + nodes[a]] += fx(f, EMC_FX_STAT);
        S[4 * NN + nodes[a]] += fx(f * energy, EMC_FX_STAT);
      }
    }
    /* particle_to_density: charge to the eight corners */
    if (node_q) {
      const double cc = charge;
      node_q[NODE(i, j, k)] += fx(cc * (1 - ry) * (1 - rx) * (1 - rz), EMC_FX_Q);
      node_q[NODE(i + 1, j, k)] += fx(cc * (1 - ry) * rx * (1 - rz), EMC_FX_Q);
      node_q[NODE(i, j, k + 1)] += fx(cc * (1 - ry) * (1 - rx) * rz, EMC_FX_Q);
      node_q[NODE(i + 1, j, k + 1)] += fx(cc * (1 - ry) * rx * rz, EMC_FX_Q);
      node_q[NODE(i, j + 1, k)] += fx(cc * ry * (1 - rx) * (1 - rz), EMC_FX_Q);
      node_q[NODE(i + 1, j + 1, k)] += fx(cc * ry * rx * (1 - rz), EMC_FX_Q);
      node_q[NODE(i, j + 1, k + 1)] += fx(cc * ry * (1 - rx) * rz, EMC_FX_Q);
      node_q[NODE(i + 1, j + 1, k + 1)] += fx(cc * ry * rx * rz, EMC_FX_Q);
    }
    /* compute_cell_charge: nearest-grid-point sums */
    if (cell_q) {
      if (pt == 0) cell_q[c] += fx(charge, EMC_FX_Q);
      else cell_q[NC + c] += fx(charge, EMC_FX_Q);
      cell_q[2 * NC + c] += fx(fabs(charge), EMC_FX_Q);
    }
  }
#undef NODE
  return 0;
}

void orc_rhs(const emc_mesh_view *m, const int64_t *node_q, double *density, double *rhs_ref, double *rhs_elim) {
  const size_t NN = (size_t)m->nnode;
  for (size_t n = 0; n < NN; ++n) {
    double d = (double)node_q[n] / EMC_FX_Q;
    if (fabs(m->pvol[n]) > 0) d = d / m->pvol[n];
    if (density) density[n] = d;
    double r = 0.0;
    if (fabs(d) > 0) r = d * m->pvol[n] + m->pdop[n];
    if (m->node_kind[n] == 1) r = m->dirichlet[n];
    if (rhs_ref) rhs_ref[n] = r;
    if (rhs_elim) rhs_elim[n] = m->node_kind[n] == 0 ? r - m->rhs_corr[n] : r;
  }
}

void orc_cell_rho(const emc_mesh_view *m, const int64_t *cell_q, double *rho) {
  const size_t NC = (size_t)m->ncell;
  for (size_t c = 0; c < NC; ++c) {
    double v = (double)cell_q[2 * NC + c] / EMC_FX_Q;
    v /= m->cvol[c];
    const double lo = 0.5 * m->da[c];
    rho[c] = v > lo ? v : lo;
  }
}

void orc_field(const emc_mesh_view *m, const double *pot, double *ex, double *ey, double *ez) {
#define P(i, j, k) pot[((size_t)(j) * m->pnx + (i)) * m->pnz + (k)]
  for (int i = 0; i < m->cnx; ++i)
    for (int k = 0; k < m->cnz; ++k)
      for (int j = 0; j < m->cny; ++j) {
        const size_t c = ((size_t)j * m->cnx + i) * m->cnz + k;
        ex[c] = -(P(i + 1, j, k) - P(i, j, k) + P(i + 1, j + 1, k) - P(i, j + 1, k) + P(i + 1, j, k + 1) -
                  P(i, j, k + 1) + P(i + 1, j + 1, k + 1) - P(i, j + 1, k + 1)) /
                (4.0 * m->dx[i]);
        ey[c] = -(P(i, j + 1, k) + P(i + 1, j + 1, k) + P(i, j + 1, k + 1) + P(i + 1, j + 1, k + 1) - P(i, j, k) -
                  P(i + 1, j, k) - P(i, j, k + 1) - P(i + 1, j, k + 1)) /
                (4.0 * m->dy[j]);
        ez[c] = -(P(i, j, k + 1) + P(i + 1, j, k + 1) + P(i, j + 1, k + 1) + P(i + 1, j + 1, k + 1) - P(i, j, k) -
                  P(i + 1, j, k) - P(i, j + 1, k) - P(i + 1, j + 1, k)) /
                (4.0 * m->dz[k]);
      }
#undef P
}

void orc_surface_rates(const emc_mesh_view *m, const double *ex, const double *ey, const double *ez,
                       const double *rho, double *e_rough, double *e_sph, double *h_rough, double *h_sph) {
  const size_t NC = (size_t)m->ncell;
  const double eps_si = m->eps_si, eps_ox = m->eps_ox;
  for (int ipt = 0; ipt < 2; ++ipt) {
    const double effmass = m->mass_eff[ipt];
    double *rr = ipt == 0 ? e_rough : h_rough, *sp = ipt == 0 ? e_sph : h_sph;
    for (size_t c = 0; c < NC; ++c) {
      const double Rho = rho[c];
      double Eeff;
      switch (m->eeff_dir[c]) {
        case 0: Eeff = fabs(ex[c]); break;
        case 1: Eeff = fabs(ey[c]); break;
        default: Eeff = fabs(ez[c]); break;
      }
      if (Eeff < 1e-20) Eeff = 1e-20;
      if (!m->flag_rough) rr[c] = 0;
      else rr[c] = 2 * PI * effmass * pow(m->surf_delta * m->surf_ail * Eeff * eps_si / (eps_si + eps_ox), 2);
      if (!m->flag_sphonon) sp[c] = 0;
      else {
        const double T0 = m->T0;
        double Wcl = pow(m->Tn, m->surf_Nbmod) * BOLTZ * T0 / (Eeff * m->field0);
        double Wq = pow(12.0 * effmass * m->em0 * (Eeff * m->field0) / m->ec0 / m->hq0 / m->hq0, -1.0 / 3.0);
        double Pf = m->surf_Pft * m->Tn + m->surf_Pfn * pow(Rho * m->conc0 / (1e17 * 1e6), m->surf_Npf) / m->Tn;
        double Wclq = (m->surf_gama * Wcl + m->surf_theta * Wq) / Pf;
        sp[c] = effmass * m->em0 * BOLTZ * T0 * m->surf_XIph * m->surf_XIph /
                    (pow(m->hq0, 3) * m->sirho * pow(9050.02, 2) * Wclq) / m->scrt0 +
                m->ec0 / (effmass * m->em0 * m->surf_Kpha * pow(m->Tn, 2.24)) / m->scrt0;
      }
    }
  }
}
