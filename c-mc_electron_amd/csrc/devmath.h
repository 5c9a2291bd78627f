// Deterministic elementary functions and the counter-based RNG of the flight.
//
// log / sin / cos follow the fdlibm methods (Sun Microsystems, freely
// distributable libm: e_log.c, k_sin.c, k_cos.c, e_rem_pio2.c), evaluated with
// IEEE double operations only.  All HIP code is compiled with
// -ffp-contract=off, so these return the same bits on the GPU and on the host.
// They stand in for glibc log/sin/cos at src/mcmodel.cpp:1944,1960,1977 and
// src/ElectronImpurityScatter.cpp:236-237 of the reference.
//
// The RNG is Philox4x32-10 (Salmon et al., SC'11) keyed per particle; it
// replaces the global ran1 stream (src/utils.cpp:413-466).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define EMC_HD __host__ __device__
// forced inlining keeps a struct whose member functions are called in registers
// instead of in per-lane scratch memory
#define EMC_INL __attribute__((always_inline))
#else
#define EMC_INL
#include <cmath>
using std::fabs;
using std::rint;
using std::sqrt;
#define EMC_HD
#endif

namespace emc {

EMC_HD inline uint32_t dm_hi(double x) { return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32); }
EMC_HD inline uint32_t dm_lo(double x) { return (uint32_t)__builtin_bit_cast(uint64_t, x); }
EMC_HD inline double dm_words(uint32_t hi, uint32_t lo) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
EMC_HD inline double dm_hex(uint64_t u) { return __builtin_bit_cast(double, u); }

EMC_HD inline double dm_log(double x) {
  const double ln2_hi = dm_hex(0x3fe62e42fee00000ull), ln2_lo = dm_hex(0x3dea39ef35793c76ull);
  const double two54 = dm_hex(0x4350000000000000ull);
  const double Lg1 = dm_hex(0x3FE5555555555593ull), Lg2 = dm_hex(0x3FD999999997FA04ull),
               Lg3 = dm_hex(0x3FD2492494229359ull), Lg4 = dm_hex(0x3FCC71C51D8E78AFull),
               Lg5 = dm_hex(0x3FC7466496CB03DEull), Lg6 = dm_hex(0x3FC39A09D078C69Full),
               Lg7 = dm_hex(0x3FC2F112DF3E5244ull);
  int32_t hx = (int32_t)dm_hi(x);
  const uint32_t lx = dm_lo(x);
  int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | (int32_t)lx) == 0) return -two54 / 0.0;
    if (hx < 0) return (x - x) / 0.0;
    k -= 54;
    x *= two54;
    hx = (int32_t)dm_hi(x);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  int32_t i = (hx + 0x95f64) & 0x100000;
  x = dm_words((uint32_t)(hx | (i ^ 0x3ff00000)), dm_lo(x));
  k += (i >> 20);
  const double f = x - 1.0;
  double dk;
  if ((0x000fffff & (2 + hx)) < 3) {
    if (f == 0.0) {
      if (k == 0) return 0.0;
      dk = (double)k;
      return dk * ln2_hi + dk * ln2_lo;
    }
    const double R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    dk = (double)k;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  const double s = f / (2.0 + f);
  dk = (double)k;
  const double z = s * s;
  i = hx - 0x6147a;
  const double w = z * z;
  const int32_t j = 0x6b851 - hx;
  const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  const double R = t2 + t1;
  if (i > 0) {
    const double hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

EMC_HD inline double dm_ksin(double x, double y, int iy) {
  const double S1 = dm_hex(0xBFC5555555555549ull), S2 = dm_hex(0x3F8111111110F8A6ull),
               S3 = dm_hex(0xBF2A01A019C161D5ull), S4 = dm_hex(0x3EC71DE357B1FE7Dull),
               S5 = dm_hex(0xBE5AE5E68A2B9CEBull), S6 = dm_hex(0x3DE5D93A5ACFD57Cull);
  const uint32_t ix = dm_hi(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return x;
  }
  const double z = x * x, v = z * x;
  const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

EMC_HD inline double dm_kcos(double x, double y) {
  const double C1 = dm_hex(0x3FA555555555554Cull), C2 = dm_hex(0xBF56C16C16C15177ull),
               C3 = dm_hex(0x3EFA01A019CB1590ull), C4 = dm_hex(0xBE927E4F809C52ADull),
               C5 = dm_hex(0x3E21EE9EBDB4B1C4ull), C6 = dm_hex(0xBDA8FAE9BE8838D4ull);
  const uint32_t ix = dm_hi(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return 1.0;
  }
  const double z = x * x;
  const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  const double qx = ix > 0x3fe90000 ? 0.28125 : dm_words(ix - 0x00200000, 0);
  const double hz = 0.5 * z - qx, a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

// Cody-Waite reduction by pi/2 (medium range |x| <= 2^19 pi/2).
EMC_HD inline int dm_rem_pio2(double x, double* y) {
  const double invpio2 = dm_hex(0x3FE45F306DC9C883ull);
  const double pio2_1 = dm_hex(0x3FF921FB54400000ull), pio2_1t = dm_hex(0x3DD0B4611A626331ull);
  const double pio2_2 = dm_hex(0x3DD0B4611A600000ull), pio2_2t = dm_hex(0x3BA3198A2E037073ull);
  const double pio2_3 = dm_hex(0x3BA3198A2E000000ull), pio2_3t = dm_hex(0x397B839A252049C1ull);
  const int32_t hx = (int32_t)dm_hi(x);
  const uint32_t ix = (uint32_t)hx & 0x7fffffff;
  double t = x < 0 ? -x : x;
  const int n = (int)(t * invpio2 + 0.5);
  const double fn = (double)n;
  double r = t - fn * pio2_1;
  double w = fn * pio2_1t;
  const int j = (int)(ix >> 20);
  y[0] = r - w;
  int i = j - (int)((dm_hi(y[0]) >> 20) & 0x7ff);
  if (i > 16) {
    t = r;
    w = fn * pio2_2;
    r = t - w;
    w = fn * pio2_2t - ((t - r) - w);
    y[0] = r - w;
    i = j - (int)((dm_hi(y[0]) >> 20) & 0x7ff);
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

EMC_HD inline double dm_sin(double x) {
  const uint32_t ix = dm_hi(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return dm_ksin(x, 0.0, 0);
  if (ix >= 0x7ff00000) return x - x;
  double y[2];
  const int n = dm_rem_pio2(x, y);
  switch (n & 3) {
    case 0: return dm_ksin(y[0], y[1], 1);
    case 1: return dm_kcos(y[0], y[1]);
    case 2: return -dm_ksin(y[0], y[1], 1);
    default: return -dm_kcos(y[0], y[1]);
  }
}

EMC_HD inline double dm_cos(double x) {
  const uint32_t ix = dm_hi(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return dm_kcos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  double y[2];
  const int n = dm_rem_pio2(x, y);
  switch (n & 3) {
    case 0: return dm_kcos(y[0], y[1]);
    case 1: return -dm_ksin(y[0], y[1], 1);
    case 2: return -dm_kcos(y[0], y[1]);
    default: return dm_ksin(y[0], y[1], 1);
  }
}

// ---------------------------------------------------------------- RNG
EMC_HD inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
  }
}

EMC_HD inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
EMC_HD inline uint64_t clone_key(uint64_t parent, uint64_t step, int ordinal) {
  return mix64(parent ^ mix64(step + 0x632BE59BD9B4E019ull * (uint64_t)(ordinal + 1)));
}
EMC_HD inline uint64_t init_key(uint64_t seed, uint64_t serial) { return mix64(mix64(seed) + serial); }
// MultipleRefresh: the draw stream of (cell, carrier type) and the key of new
// particle i of that cell and type, copied from `parent`
EMC_HD inline uint64_t mr_stream_key(long long cell, int type) {
  return mix64(0x4D554C5449524546ull ^ mix64((uint64_t)(2 * cell + type) + 1u));
}
EMC_HD inline uint64_t mr_child_key(uint64_t parent, uint64_t step, long long i) {
  return mix64(parent ^ mix64(step + 0xD1B54A32D192ED03ull * (uint64_t)(i + 1)));
}

// One particle's stream for one step: draw n is element n%2 of Philox block n/2;
// 52-bit mantissa -> (m + 0.5) 2^-52, strictly inside (0,1).
// A zero defined per lane (in a vector register).  The draw counters below start
// from it: they change only through control flow (how many draws a particle
// makes), and a value that starts per lane can never be merged into one
// wave-wide register, whatever the compiler concludes about the branches.
EMC_HD inline uint32_t lane_zero() {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
#else
  return 0u;
#endif
}

struct Rng {
  uint32_t k0, k1, c1, c2, block;
  int have;
  double spare;
  EMC_HD void init(uint64_t key, uint64_t step) {
    k0 = (uint32_t)key; k1 = (uint32_t)(key >> 32);
    c1 = (uint32_t)step; c2 = (uint32_t)(step >> 32);
    const uint32_t z = lane_zero();
    block = z; have = (int)z; spare = 0.0;
  }
  EMC_HD double next() {
    if (have) { have = 0; return spare; }
    uint32_t c[4] = {block, c1, c2, 0u};
    philox4x32_10(c, k0, k1);
    ++block;
    const uint64_t a = (((uint64_t)c[1] << 32) | c[0]) >> 12;
    const uint64_t b = (((uint64_t)c[3] << 32) | c[2]) >> 12;
    spare = ((double)b + 0.5) * 2.220446049250313080847263336181640625e-16;
    have = 1;
    return ((double)a + 0.5) * 2.220446049250313080847263336181640625e-16;
  }
};

}  // namespace emc
