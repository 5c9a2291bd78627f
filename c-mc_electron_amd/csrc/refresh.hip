// MultipleRefresh (src/MultipleRefresh.cpp:9-143) on the GPU: per-cell,
// per-carrier-type resampling of the population to the desired counts.
//
//   k_mr_keys     group (2*cell + type) and key of every slot
//   radix sorts   by key, then (stable) by group: each group's particles in key order
//   k_mr_bounds   first and last position of each group in the sorted order; live count
//   k_mr_plan     per group: charge sums, the keep-or-resample test, output count
//   scan          output offsets
//   k_mr_write    per group: the kept particles, or new_num copies drawn with
//                 probability |q| / sum|q|, charge sum(q) / new_num
//
// One thread per group walks its particles in key order, so every sum is taken
// in the same order as the oracle's and the output is identical slot for slot.
// The radix sorts and the scan are hipCUB's.  Runs every mr_step steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "devmath.h"
#include "emc_const.h"
#include "kernels.h"

namespace emc {

constexpr double MR_RPAR = 2.0, MR_RSQR = 2.0, MR_PCKILL = 1.0e-30;  // src/mcmodel.cpp:3883-3885

__device__ inline long long mr_cell(const MrArgs& a, uint32_t pk) {
  return ((long long)pk_j(pk) * a.cnx + pk_i(pk)) * a.cnz + pk_k(pk);
}

__global__ __launch_bounds__(256) void k_mr_keys(MrArgs a) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < a.n; q += (long long)gridDim.x * blockDim.x) {
    const uint32_t pk = a.P.pk[q];
    const bool on = pk != PK_DEAD;
    a.gk[q] = on ? (unsigned)(2 * mr_cell(a, pk) + pk_type(pk)) : (unsigned)a.ngroup;
    a.sk[q] = a.P.key[q];
    a.idx[q] = (unsigned)q;
  }
}

// group key of each position after the sort by particle key
__global__ __launch_bounds__(256) void k_mr_gather_group(MrArgs a) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += (long long)gridDim.x * blockDim.x)
    a.gk2[p] = a.gk[a.idx2[p]];
}

__global__ __launch_bounds__(256) void k_mr_clear(MrArgs a) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroup; g += (long long)gridDim.x * blockDim.x) {
    a.gstart[g] = 0;
    a.gend[g] = 0;
  }
}

// after the sorts: gk (group) and idx (slot) in (group, key) order; empty slots
// sort last, so the live count is the position of the first of them
__global__ __launch_bounds__(256) void k_mr_bounds(MrArgs a) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += (long long)gridDim.x * blockDim.x) {
    const unsigned g = a.gk[p];
    if (g >= (unsigned)a.ngroup) {
      if (p == 0 || a.gk[p - 1] < (unsigned)a.ngroup) *a.nlive = (unsigned long long)p;
      continue;
    }
    if (p == a.n - 1) *a.nlive = (unsigned long long)a.n;
    if (p == 0 || a.gk[p - 1] != g) a.gstart[g] = (unsigned)p;
    if (p == a.n - 1 || a.gk[p + 1] != g) a.gend[g] = (unsigned)(p + 1);
  }
}

// gen_new_par_mr's test (src/MultipleRefresh.cpp:17-36): V = the group without
// |q| < pckill; keep V when Rpar bounds its size and the charge spread is small
__global__ __launch_bounds__(256) void k_mr_plan(MrArgs a) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroup; g += (long long)gridDim.x * blockDim.x) {
    const unsigned p0 = a.gstart[g], p1 = a.gend[g];
    long long nv = 0;
    double tot = 0, tmp = 0;
    for (unsigned p = p0; p < p1; ++p) {
      const double q = a.P.charge[a.idx[p]];
      if (fabs(q) < MR_PCKILL) continue;
      ++nv;
      tot += q;
      tmp += q * q;
    }
    const long long c = g >> 1;
    const int new_num = (g & 1) ? a.desired_h[c] : a.desired_e[c];
    const bool keep = ((double)nv > new_num / MR_RPAR) && ((double)nv < new_num * MR_RPAR) &&
                      (tmp < tot * tot * MR_RSQR / (double)nv);
    a.cnt[g] = nv == 0 ? 0ull : (unsigned long long)(keep ? nv : (long long)(new_num > 0 ? new_num : 0));
    a.keep[g] = keep ? 1 : 0;
    a.tot[g] = tot;
  }
}

__device__ inline void mr_copy(const PartDev& P, long long s, const PartDev& Q, long long d) {
  Q.x[d] = P.x[s]; Q.y[d] = P.y[s]; Q.z[d] = P.z[s];
  Q.kx[d] = P.kx[s]; Q.ky[d] = P.ky[s]; Q.kz[d] = P.kz[s];
  Q.energy[d] = P.energy[s]; Q.charge[d] = P.charge[s];
  Q.pk[d] = P.pk[s]; Q.itet[d] = P.itet[s]; Q.isym[d] = P.isym[s]; Q.key[d] = P.key[s];
}

__global__ __launch_bounds__(256) void k_mr_write(MrArgs a) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroup; g += (long long)gridDim.x * blockDim.x) {
    const unsigned long long m = a.cnt[g];
    if (m == 0) continue;
    const long long d0 = (long long)a.off[g];
    const unsigned p0 = a.gstart[g], p1 = a.gend[g];
    if (a.keep[g]) {
      long long r = 0;
      for (unsigned p = p0; p < p1; ++p) {
        const long long s = a.idx[p];
        if (fabs(a.P.charge[s]) < MR_PCKILL) continue;
        mr_copy(a.P, s, a.Q, d0 + r++);
      }
      continue;
    }
    const long long c = g >> 1;
    const int t = (int)(g & 1);
    const double tot = a.tot[g];
    const double aver = tot / (double)(long long)m;
    Rng rng;
    rng.init(mr_stream_key(c, t), a.step);
    for (long long i = 0; i < (long long)m; ++i) {
      // lower_bound over the running sums of |q| in key order: the first
      // particle whose running sum reaches the draw (the last one otherwise)
      const double sv = rng.next() * fabs(tot);
      double cum = 0;
      long long parent = -1;
      for (unsigned p = p0; p < p1; ++p) {
        const long long s = a.idx[p];
        const double q = a.P.charge[s];
        if (fabs(q) < MR_PCKILL) continue;
        cum += fabs(q);
        parent = s;
        if (cum >= sv) break;
      }
      mr_copy(a.P, parent, a.Q, d0 + i);
      a.Q.charge[d0 + i] = aver;
      a.Q.key[d0 + i] = mr_child_key(a.P.key[parent], a.step, i);
    }
  }
}

}  // namespace emc

using namespace emc;

static int mr_grid(long long n) {
  long long g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

static int bits_for(long long v) {
  int b = 1;
  while (b < 32 && (1ll << b) <= v) ++b;
  return b;
}

// Temporary storage hipCUB needs for `n` slots and `ngroup` groups.
size_t emc_mr_temp_bytes(long long n, long long ngroup) {
  size_t a = 0, b = 0, c = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                           (unsigned*)nullptr, (unsigned*)nullptr, (int)n);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (unsigned*)nullptr, (unsigned*)nullptr, (unsigned*)nullptr,
                                           (unsigned*)nullptr, (int)n, 0, bits_for(ngroup));
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                         (int)ngroup);
  return std::max(a, std::max(b, c));
}

// Everything up to the output offsets and the live count; the output count is
// off[ngroup-1] + cnt[ngroup-1].
hipError_t emc_launch_mr_plan(const MrArgs& a, hipStream_t s) {
  const int g = mr_grid(a.n), gg = mr_grid(a.ngroup);
  hipLaunchKernelGGL(k_mr_keys, dim3(g), dim3(256), 0, s, a);
  size_t tb = a.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.sk, a.sk2, a.idx, a.idx2, (int)a.n, 0, 64, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mr_gather_group, dim3(g), dim3(256), 0, s, a);
  tb = a.temp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.gk2, a.gk, a.idx2, a.idx, (int)a.n, 0, bits_for(a.ngroup), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mr_clear, dim3(gg), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_mr_bounds, dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_mr_plan, dim3(gg), dim3(256), 0, s, a);
  tb = a.temp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(a.temp, tb, a.cnt, a.off, (int)a.ngroup, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t emc_launch_mr_write(const MrArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mr_write, dim3(mr_grid(a.ngroup)), dim3(256), 0, s, a);
  return hipGetLastError();
}
