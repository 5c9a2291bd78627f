/*
 * omath.c - deterministic elementary functions and the counter-based RNG of
 * the oracle (test infrastructure only; see oracle.h).
 *
 * log, sin and cos restate the published fdlibm methods (Sun Microsystems
 * freely distributable libm, e_log.c / k_sin.c / k_cos.c / e_rem_pio2.c):
 * table-free argument reduction plus minimax polynomials, evaluated with
 * IEEE double operations only (no FMA contraction), so the same sequence on
 * the GPU gives bit-identical results.  They replace glibc's log/sin/cos used
 * at src/mcmodel.cpp:1944,1960,1977 and src/ElectronImpurityScatter.cpp:236-237
 * (error < 1 ulp; checked against glibc in tests/test_oracle_math.py).
 *
 * Philox4x32-10 follows Salmon et al., "Parallel random numbers: as easy as
 * 1, 2, 3" (SC'11); it replaces ran1 (src/utils.cpp:413-466).
 */
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static inline uint32_t hi_word(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return (uint32_t)(u >> 32);
}
static inline uint32_t lo_word(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return (uint32_t)u;
}
static inline double from_words(uint32_t hi, uint32_t lo) {
  uint64_t u = ((uint64_t)hi << 32) | lo;
  double x;
  memcpy(&x, &u, 8);
  return x;
}
static inline double hexd(uint64_t u) {
  double x;
  memcpy(&x, &u, 8);
  return x;
}

/* ---------------------------------------------------------------- log */
double orc_log(double x) {
  const double ln2_hi = hexd(0x3fe62e42fee00000ull), ln2_lo = hexd(0x3dea39ef35793c76ull);
  const double two54 = hexd(0x4350000000000000ull);
  const double Lg1 = hexd(0x3FE5555555555593ull), Lg2 = hexd(0x3FD999999997FA04ull),
               Lg3 = hexd(0x3FD2492494229359ull), Lg4 = hexd(0x3FCC71C51D8E78AFull),
               Lg5 = hexd(0x3FC7466496CB03DEull), Lg6 = hexd(0x3FC39A09D078C69Full),
               Lg7 = hexd(0x3FC2F112DF3E5244ull);
  int32_t hx = (int32_t)hi_word(x);
  uint32_t lx = lo_word(x);
  int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | lx) == 0) return -two54 / 0.0;
    if (hx < 0) return (x - x) / 0.0;
    k -= 54;
    x *= two54;
    hx = (int32_t)hi_word(x);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  int32_t i = (hx + 0x95f64) & 0x100000;
  x = from_words((uint32_t)(hx | (i ^ 0x3ff00000)), lo_word(x));
  k += (i >> 20);
  double f = x - 1.0;
  double dk;
  if ((0x000fffff & (2 + hx)) < 3) {
    if (f == 0.0) {
      if (k == 0) return 0.0;
      dk = (double)k;
      return dk * ln2_hi + dk * ln2_lo;
    }
    double R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    dk = (double)k;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  double s = f / (2.0 + f);
  dk = (double)k;
  double z = s * s;
  i = hx - 0x6147a;
  double w = z * z;
  int32_t j = 0x6b851 - hx;
  double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  double R = t2 + t1;
  if (i > 0) {
    double hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

/* ---------------------------------------------------------------- sin / cos */
static double k_sin(double x, double y, int iy) {
  const double S1 = hexd(0xBFC5555555555549ull), S2 = hexd(0x3F8111111110F8A6ull),
               S3 = hexd(0xBF2A01A019C161D5ull), S4 = hexd(0x3EC71DE357B1FE7Dull),
               S5 = hexd(0xBE5AE5E68A2B9CEBull), S6 = hexd(0x3DE5D93A5ACFD57Cull);
  uint32_t ix = hi_word(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return x;
  }
  double z = x * x;
  double v = z * x;
  double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

static double k_cos(double x, double y) {
  const double C1 = hexd(0x3FA555555555554Cull), C2 = hexd(0xBF56C16C16C15177ull),
               C3 = hexd(0x3EFA01A019CB1590ull), C4 = hexd(0xBE927E4F809C52ADull),
               C5 = hexd(0x3E21EE9EBDB4B1C4ull), C6 = hexd(0xBDA8FAE9BE8838D4ull);
  uint32_t ix = hi_word(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return 1.0;
  }
  double z = x * x;
  double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  double qx = ix > 0x3fe90000 ? 0.28125 : from_words(ix - 0x00200000, 0);
  double hz = 0.5 * z - qx;
  double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

/* Cody-Waite reduction by pi/2 for |x| <= 2^19 pi/2 (fdlibm e_rem_pio2 medium
 * case); larger arguments are not produced on this path. */
static int rem_pio2(double x, double *y) {
  const double invpio2 = hexd(0x3FE45F306DC9C883ull);
  const double pio2_1 = hexd(0x3FF921FB54400000ull), pio2_1t = hexd(0x3DD0B4611A626331ull);
  const double pio2_2 = hexd(0x3DD0B4611A600000ull), pio2_2t = hexd(0x3BA3198A2E037073ull);
  const double pio2_3 = hexd(0x3BA3198A2E000000ull), pio2_3t = hexd(0x397B839A252049C1ull);
  int32_t hx = (int32_t)hi_word(x);
  uint32_t ix = (uint32_t)hx & 0x7fffffff;
  double t = x < 0 ? -x : x;
  int n = (int)(t * invpio2 + 0.5);
  double fn = (double)n;
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t;
  int j = (int)(ix >> 20);
  y[0] = r - w;
  int i = j - (int)((hi_word(y[0]) >> 20) & 0x7ff);
  if (i > 16) {
    t = r;
    w = fn * pio2_2;
    r = t - w;
    w = fn * pio2_2t - ((t - r) - w);
    y[0] = r - w;
    i = j - (int)((hi_word(y[0]) >> 20) & 0x7ff);
    if (i > 49) {
      t = r;
      w = fn * pio2_3;
      r = t - w;
      w = fn * pio2_3t - ((t - r) - w);
      y[0] = r - w;
    }
  }
  y[1] = (r - y[0]) - w;
  if (hx < 0) {
    y[0] = -y[0];
    y[1] = -y[1];
    return -n;
  }
  return n;
}

double orc_sin(double x) {
  uint32_t ix = hi_word(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return k_sin(x, 0.0, 0);
  if (ix >= 0x7ff00000) return x - x;
  double y[2];
  int n = rem_pio2(x, y);
  switch (n & 3) {
    case 0: return k_sin(y[0], y[1], 1);
    case 1: return k_cos(y[0], y[1]);
    case 2: return -k_sin(y[0], y[1], 1);
    default: return -k_cos(y[0], y[1]);
  }
}

double orc_cos(double x) {
  uint32_t ix = hi_word(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return k_cos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  double y[2];
  int n = rem_pio2(x, y);
  switch (n & 3) {
    case 0: return k_cos(y[0], y[1]);
    case 1: return -k_sin(y[0], y[1], 1);
    case 2: return -k_cos(y[0], y[1]);
    default: return k_sin(y[0], y[1], 1);
  }
}

/* ---------------------------------------------------------------- Philox4x32-10 */
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint64_t orc_mix64(uint64_t z) {
  /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_clone_key(uint64_t parent, uint64_t step, int ordinal) {
  return orc_mix64(parent ^ orc_mix64(step + 0x632BE59BD9B4E019ull * (uint64_t)(ordinal + 1)));
}

uint64_t orc_init_key(uint64_t seed, uint64_t serial) { return orc_mix64(orc_mix64(seed) + serial); }

/* 52 random bits -> (x + 0.5) * 2^-52, exactly representable, in (0,1) */
static inline double u52(uint32_t lo, uint32_t hi) {
  uint64_t b = (((uint64_t)hi << 32) | lo) >> 12;
  return ((double)b + 0.5) * 2.220446049250313080847263336181640625e-16;
}

void orc_uniforms(uint64_t key, uint64_t step, int n, double *out) {
  uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
  for (int i = 0; i < n; i += 2) {
    uint32_t c[4] = {(uint32_t)(i / 2), (uint32_t)step, (uint32_t)(step >> 32), 0u}, o[4];
    orc_philox4x32(c, k, o);
    out[i] = u52(o[0], o[1]);
    if (i + 1 < n) out[i + 1] = u52(o[2], o[3]);
  }
}
