/*
 * orng.h - the counter-based draw stream of the oracle (test infrastructure
 * only): Philox4x32-10 with key = 64-bit stream key and counter = (draw block,
 * step lo, step hi, 0); each block gives two draws ((bits >> 12) + 0.5) * 2^-52
 * from its two 64-bit halves, low half first.
 */
#ifndef EMC_ORACLE_RNG_H
#define EMC_ORACLE_RNG_H

#include <stdint.h>

#include "oracle.h"

typedef struct {
  uint32_t k0, k1, c1, c2;
  uint32_t block;
  int have;
  double spare;
} Rng;

static inline double rng_next(Rng *r) {
  if (r->have) {
    r->have = 0;
    return r->spare;
  }
  uint32_t c[4] = {r->block, r->c1, r->c2, 0u}, k[2] = {r->k0, r->k1}, o[4];
  orc_philox4x32(c, k, o);
  r->block++;
  uint64_t a = (((uint64_t)o[1] << 32) | o[0]) >> 12, b = (((uint64_t)o[3] << 32) | o[2]) >> 12;
  r->spare = ((double)b + 0.5) * 2.220446049250313080847263336181640625e-16;
  r->have = 1;
  return ((double)a + 0.5) * 2.220446049250313080847263336181640625e-16;
}

static inline void rng_init(Rng *r, uint64_t key, uint64_t step) {
  r->k0 = (uint32_t)key;
  r->k1 = (uint32_t)(key >> 32);
  r->c1 = (uint32_t)step;
  r->c2 = (uint32_t)(step >> 32);
  r->block = 0;
  r->have = 0;
  r->spare = 0;
}

#endif
