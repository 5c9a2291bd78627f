/*
 * omr.c - CPU restatement of MultipleRefresh (src/MultipleRefresh.cpp:9-143),
 * test infrastructure only.
 *
 * Per cell and carrier type (src/MultipleRefresh.cpp:79-135): particles with
 * |q| < pckill are dropped; the rest (V, in ascending key order, the defined
 * order of a cell's list) are resampled by gen_new_par_mr
 * (src/MultipleRefresh.cpp:9-59) to the cell's desired count, each copy drawn
 * with probability |q| / sum|q| and given the charge sum(q) / new_num, unless
 * the count is within a factor Rpar of the desired one and the charge spread
 * is small (sum q^2 < (sum q)^2 Rsqr / n); Rpar = Rsqr = 2, pckill = 1e-30
 * (src/mcmodel.cpp:3883-3885).
 *
 * Defined where the reference uses its global ran1 stream and list order:
 *   - the draws of (cell c, type t) at step s are the stream of key
 *     orc_mr_stream_key(c, t) and step s (orng.h), one draw per new particle;
 *   - new particle i of a refreshed (c, t) copies its parent with key
 *     orc_mr_child_key(parent key, s, i);
 *   - the output runs over cells in linear order (j*cnx + i)*cnz + k, electrons
 *     before holes, kept particles in key order, new ones in draw order;
 *   - a total charge of 0 (every draw would be 0, where the reference stops
 *     with a warning) takes the first particle as parent.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orng.h"

#define MR_RPAR 2.0
#define MR_RSQR 2.0
#define MR_PCKILL 1.0e-30

uint64_t orc_mr_stream_key(int64_t cell, int type) {
  return orc_mix64(0x4D554C5449524546ull ^ orc_mix64((uint64_t)(2 * cell + type) + 1u));
}

uint64_t orc_mr_child_key(uint64_t parent, uint64_t step, int64_t i) {
  return orc_mix64(parent ^ orc_mix64(step + 0xD1B54A32D192ED03ull * (uint64_t)(i + 1)));
}

typedef struct {
  int64_t g;  /* 2*cell + type */
  uint64_t key;
  int64_t slot;
} Ent;

static int ent_cmp(const void *a, const void *b) {
  const Ent *x = (const Ent *)a, *y = (const Ent *)b;
  if (x->g != y->g) return x->g < y->g ? -1 : 1;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return 0;
}

static void copy_particle(const emc_particles *in, int64_t s, emc_particles *out, int64_t d) {
  out->x[d] = in->x[s]; out->y[d] = in->y[s]; out->z[d] = in->z[s];
  out->kx[d] = in->kx[s]; out->ky[d] = in->ky[s]; out->kz[d] = in->kz[s];
  out->energy[d] = in->energy[s]; out->charge[d] = in->charge[s];
  out->i[d] = in->i[s]; out->j[d] = in->j[s]; out->k[d] = in->k[s];
  out->itet[d] = in->itet[s]; out->isym[d] = in->isym[s]; out->ptype[d] = in->ptype[s];
  out->key[d] = in->key[s];
}

int64_t orc_multiple_refresh(const emc_mesh_view *m, const emc_particles *in, int64_t n, const int32_t *desired_e,
                             const int32_t *desired_h, uint64_t step, emc_particles *out, int64_t cap) {
  Ent *e = (Ent *)malloc(sizeof(Ent) * (size_t)(n > 0 ? n : 1));
  int64_t ne = 0;
  for (int64_t s = 0; s < n; ++s) {
    if (in->i[s] < 0) continue;
    const int64_t c = ((int64_t)in->j[s] * m->cnx + in->i[s]) * m->cnz + in->k[s];
    e[ne].g = 2 * c + in->ptype[s];
    e[ne].key = in->key[s];
    e[ne].slot = s;
    ++ne;
  }
  qsort(e, (size_t)ne, sizeof(Ent), ent_cmp);
  int64_t *v = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ne > 0 ? ne : 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)(ne > 0 ? ne : 1));
  int64_t nout = 0;
  for (int64_t a = 0; a < ne;) {
    int64_t b = a;
    while (b < ne && e[b].g == e[a].g) ++b;
    const int64_t c = e[a].g / 2;
    const int t = (int)(e[a].g % 2);
    /* V: the group without the killed particles */
    int64_t nv = 0;
    for (int64_t q = a; q < b; ++q)
      if (fabs(in->charge[e[q].slot]) >= MR_PCKILL) v[nv++] = e[q].slot;
    a = b;
    if (nv == 0) continue; /* gen_new_par_mr returns false; nothing to keep */
    const int new_num = t == 0 ? desired_e[c] : desired_h[c];
    double s = 0, tot = 0, tmp = 0;
    for (int64_t q = 0; q < nv; ++q) {
      const double ch = in->charge[v[q]];
      s += fabs(ch);
      w[q] = s;
      tot += ch;
      tmp += ch * ch;
    }
    const int keep = ((double)nv > new_num / MR_RPAR) && ((double)nv < new_num * MR_RPAR) &&
                     (tmp < tot * tot * MR_RSQR / (double)nv);
    if (keep) {
      for (int64_t q = 0; q < nv; ++q) {
        if (nout >= cap) goto overflow;
        copy_particle(in, v[q], out, nout++);
      }
      continue;
    }
    Rng r;
    rng_init(&r, orc_mr_stream_key(c, t), step);
    const double aver = tot / new_num;
    for (int q = 0; q < new_num; ++q) {
      const double sv = rng_next(&r) * fabs(tot);
      int64_t idx = 0; /* first weight >= sv (lower_bound over 0, w[0], w[1], ...) */
      while (idx < nv - 1 && w[idx] < sv) ++idx;
      if (nout >= cap) goto overflow;
      copy_particle(in, v[idx], out, nout);
      out->charge[nout] = aver;
      out->key[nout] = orc_mr_child_key(in->key[v[idx]], step, q);
      ++nout;
    }
  }
  free(e);
  free(v);
  free(w);
  return nout;
overflow:
  free(e);
  free(v);
  free(w);
  return -1;
}
