// HIP kernels of the MI355X ensemble Monte Carlo step (gfx950, wave64).
//
//   k_fly_fast   one lane per particle: drift of those with no event this step,
//                the rest listed for
//   k_fly_slow   one lane per listed particle, full-band event loop (fly_core.h)
//   k_fly_clones same for GENERATE clones of this step
//   k_merge      surviving clones -> holes left by caught particles, then append
//   k_deposit    CIC charge, |q|-weighted node statistics, NGP cell sums (int64 atomics)
//   k_rhs        node density and Poisson right-hand side (eliminated system)
//   k_field      cell fields from node potential, Rho clamp, surface rates
//   k_pcg_*      Jacobi-preconditioned CG on the 7-point operator
//
// All HBM-bound integer/FP64 work: no MFMA.  Compiled with -ffp-contract=off so
// the FP64 results match the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include "fly_core.h"
#include "kernels.h"

namespace emc {

// ------------------------------------------------------------------ reductions
template <class T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// per-thread counters folded once per wave
__device__ inline void flush_counters(const TallyDev& T, long long ev, long long ph, long long im, long long ca) {
  ev = wave_sum(ev);
  ph = wave_sum(ph);
  im = wave_sum(im);
  ca = wave_sum(ca);
  if ((threadIdx.x & 63) == 0) {
    if (ev) atomicAdd(&T.counters[CNT_EVENTS], (unsigned long long)ev);
    if (ph) atomicAdd(&T.counters[CNT_PHONON], (unsigned long long)ph);
    if (im) atomicAdd(&T.counters[CNT_IMPURITY], (unsigned long long)im);
    if (ca) atomicAdd(&T.counters[CNT_CATCH], (unsigned long long)ca);
  }
}

__device__ inline void finish_drop(const TallyDev& T, const Flight& f) {
  if (f.drop_reason >= 0) atomicAdd(&T.counters[CNT_DROP0 + f.drop_reason], 1ull);
}

// ------------------------------------------------------------------ flight
__device__ inline void load_particle(Flight& f, const FlyArgs& a, long long q, uint32_t pk) {
  f.clones = a.clones;
  f.nclones = a.nclones;
  f.clone_cap = a.clone_cap;
  f.x = a.P.x[q]; f.y = a.P.y[q]; f.z = a.P.z[q];
  f.kx = a.P.kx[q]; f.ky = a.P.ky[q]; f.kz = a.P.kz[q];
  f.energy = a.P.energy[q]; f.charge = a.P.charge[q];
  f.icell = pk_i(pk); f.jcell = pk_j(pk); f.kcell = pk_k(pk); f.par_type = pk_type(pk);
  f.itet = a.P.itet[q]; f.isym = a.P.isym[q];
  f.key = a.P.key[q]; f.step = a.step;
  f.left_time = a.M.dt;
  f.nclone = 0; f.flag_heat = a.flag_heat;
  f.n_events = 0; f.n_phonon = 0; f.n_impurity = 0;
  f.rng.init(f.key, a.step);
}

// Fast pass: every particle whose first event is the end of the time step drifts
// and is written back (position, wave vector and energy; cell, tetrahedron and
// wedge are unchanged).  The others are listed, wave-compacted, for k_fly_slow,
// which restarts them from the untouched state: the RNG is counter based, so the
// full event loop draws the same numbers and the result equals one pass of fly().
__global__ __launch_bounds__(256) void k_fly_fast(FlyArgs a) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long ev = 0;
  const int lane = threadIdx.x & 63;
  // the loop bound is uniform per block so every lane reaches the ballot
  for (long long base = (long long)blockIdx.x * blockDim.x; base < a.n; base += stride) {
    const long long q = base + threadIdx.x;
    bool slow = false;
    if (q < a.n) {
      const uint32_t pk = a.P.pk[q];
      if (pk != PK_DEAD) {
        Flight f(a.B, a.M, a.F, a.T);
        load_particle(f, a, q, pk);
        if (f.fly_fast()) {
          ev += f.n_events;
          a.P.x[q] = f.x; a.P.y[q] = f.y; a.P.z[q] = f.z;
          a.P.kx[q] = f.kx; a.P.ky[q] = f.ky; a.P.kz[q] = f.kz;
          a.P.energy[q] = f.energy;
        } else
          slow = true;
      }
    }
    const unsigned long long m = __ballot(slow);
    if (m) {
      const int leader = __builtin_ctzll(m);
      unsigned long long b0 = 0;
      if (lane == leader) b0 = atomicAdd(a.nslow, (unsigned long long)__popcll(m));
      b0 = __shfl(b0, leader, 64);
      if (slow) a.slow[b0 + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)q;
    }
  }
  ev = wave_sum(ev);
  if (lane == 0 && ev) atomicAdd(&a.T.counters[CNT_EVENTS], (unsigned long long)ev);
}

// Full event loop for the particles k_fly_fast listed; one lane per entry.
__global__ __launch_bounds__(256) void k_fly_slow(FlyArgs a) {
  long long ev = 0, ph = 0, im = 0, ca = 0;
  const long long ns = (long long)*a.nslow;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ns; e += stride) {
    const long long q = a.slow[e];
    const uint32_t pk = a.P.pk[q];
    Flight f(a.B, a.M, a.F, a.T);
    load_particle(f, a, q, pk);
    f.fly();
    ev += f.n_events; ph += f.n_phonon; im += f.n_impurity;
    if (f.Flag_Catch) {
      ++ca;
      finish_drop(a.T, f);
      a.P.pk[q] = PK_DEAD;
      const unsigned long long d = atomicAdd(a.ndead, 1ull);
      if (d < a.dead_cap) a.dead[d] = q;
      continue;
    }
    a.P.x[q] = f.x; a.P.y[q] = f.y; a.P.z[q] = f.z;
    a.P.kx[q] = f.kx; a.P.ky[q] = f.ky; a.P.kz[q] = f.kz;
    a.P.energy[q] = f.energy;
    a.P.pk[q] = pk_make(f.par_type, f.icell, f.jcell, f.kcell);
    a.P.itet[q] = f.itet; a.P.isym[q] = f.isym;
  }
  flush_counters(a.T, ev, ph, im, ca);
}

// Clones [begin, end) of the clone buffer fly the remaining time of the step.
__global__ __launch_bounds__(256) void k_fly_clones(FlyArgs a, unsigned long long begin, unsigned long long end) {
  long long ev = 0, ph = 0, im = 0, ca = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)begin + blockIdx.x * blockDim.x + threadIdx.x; q < (long long)end; q += stride) {
    CloneRec& c = a.clones[q];
    Flight f(a.B, a.M, a.F, a.T);
    f.clones = a.clones;
    f.nclones = a.nclones;
    f.clone_cap = a.clone_cap;
    f.x = c.x; f.y = c.y; f.z = c.z; f.kx = c.kx; f.ky = c.ky; f.kz = c.kz;
    f.energy = c.energy; f.charge = c.charge; f.left_time = c.left_time;
    f.icell = pk_i(c.pk); f.jcell = pk_j(c.pk); f.kcell = pk_k(c.pk); f.par_type = pk_type(c.pk);
    f.itet = c.itet; f.isym = c.isym; f.key = c.key; f.step = a.step;
    f.nclone = 0; f.flag_heat = a.flag_heat;
    f.n_events = 0; f.n_phonon = 0; f.n_impurity = 0;
    f.rng.init(f.key, a.step);
    f.fly();
    ev += f.n_events; ph += f.n_phonon; im += f.n_impurity;
    if (f.Flag_Catch) {
      ++ca;
      finish_drop(a.T, f);
      c.pk = PK_DEAD;
      continue;
    }
    c.x = f.x; c.y = f.y; c.z = f.z; c.kx = f.kx; c.ky = f.ky; c.kz = f.kz;
    c.energy = f.energy;
    c.pk = pk_make(f.par_type, f.icell, f.jcell, f.kcell);
    c.itet = f.itet; c.isym = f.isym;
  }
  flush_counters(a.T, ev, ph, im, ca);
}

// Surviving clones take the slots of caught particles first, then go past n.
__global__ __launch_bounds__(256) void k_merge(FlyArgs a, unsigned long long nclones, unsigned long long ndead,
                                               unsigned long long* ticket) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < (long long)nclones; q += stride) {
    const CloneRec& c = a.clones[q];
    if (c.pk == PK_DEAD) continue;
    const unsigned long long m = atomicAdd(ticket, 1ull);
    long long slot;
    if (m < ndead && m < a.dead_cap) slot = a.dead[m];
    else slot = a.n + (long long)(m - (ndead < a.dead_cap ? ndead : a.dead_cap));
    if (slot >= a.cap) {
      atomicAdd(&a.T.counters[CNT_DROP0 + 8], 1ull);
      continue;
    }
    a.P.x[slot] = c.x; a.P.y[slot] = c.y; a.P.z[slot] = c.z;
    a.P.kx[slot] = c.kx; a.P.ky[slot] = c.ky; a.P.kz[slot] = c.kz;
    a.P.energy[slot] = c.energy; a.P.charge[slot] = c.charge;
    a.P.pk[slot] = c.pk; a.P.itet[slot] = c.itet; a.P.isym[slot] = c.isym; a.P.key[slot] = c.key;
  }
}

// ------------------------------------------------------------------ deposit
// statistic (src/mcmodel.cpp:1167-1408), particle_to_density (:348-417) and
// compute_cell_charge (:2664-2712) in one pass over the particles.
__global__ __launch_bounds__(256) void k_deposit(DepositArgs a) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const size_t NN = a.nnode, NC = a.ncell;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < a.n; q += stride) {
    const uint32_t pk = a.P.pk[q];
    if (pk == PK_DEAD) continue;
    const int i = pk_i(pk), j = pk_j(pk), k = pk_k(pk), pt = pk_type(pk);
    const double x = a.P.x[q], y = a.P.y[q], z = a.P.z[q], charge = a.P.charge[q], energy = a.P.energy[q];
    const double rx = (x - a.M.lx[i]) / a.M.dx[i];
    const double ry = (y - a.M.ly[j]) / a.M.dy[j];
    const double rz = (z - a.M.lz[k]) / a.M.dz[k];
    const size_t c = ((size_t)j * a.M.cnx + i) * a.M.cnz + k;
    const int pnx = a.M.pnx, pnz = a.M.pnz;
    auto node = [&](int ii, int jj, int kk) { return ((size_t)jj * pnx + ii) * pnz + kk; };
    const size_t nd[8] = {node(i, j, k),         node(i + 1, j, k),         node(i, j, k + 1),
                          node(i + 1, j, k + 1), node(i, j + 1, k),         node(i + 1, j + 1, k),
                          node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1)};
    if (a.cell_n) atomicAdd(&a.cell_n[c], 1ull);
    if (a.node_stat) {
      const int isym = a.P.isym[q];
      const double* vt = a.B.tr[a.P.itet[q]].vt;
      const int* ms = a.B.matsym;
      const double v3[3] = {vt[0], vt[1], vt[2]};
      const double vx = v3[ms[0 * 48 + isym] - 1] * ms[3 * 48 + isym];
      const double vy = v3[ms[1 * 48 + isym] - 1] * ms[4 * 48 + isym];
      const double vz = v3[ms[2 * 48 + isym] - 1] * ms[5 * 48 + isym];
      unsigned long long* S = a.node_stat + (size_t)(pt == 0 ? 0 : 5) * NN;
      const double cc[8] = {charge * (1 - rx) * (1 - ry) * (1 - rz), charge * rx * (1 - ry) * (1 - rz),
                            charge * (1 - rx) * (1 - ry) * rz,       charge * rx * (1 - ry) * rz,
                            charge * (1 - rx) * ry * (1 - rz),       charge * rx * ry * (1 - rz),
                            charge * (1 - rx) * ry * rz,             charge * rx * ry * rz};
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const double f = fabs(cc[e]);
        atomicAdd(&S[0 * NN + nd[e]], (unsigned long long)fxd(vx * f, 4294967296.0));
        atomicAdd(&S[1 * NN + nd[e]], (unsigned long long)fxd(vy * f, 4294967296.0));
        atomicAdd(&S[2 * NN + nd[e]], (unsigned long long)fxd(vz * f, 4294967296.0));
        atomicAdd(&S[3 * NN + nd[e]], (unsigned long long)fxd(f, 4294967296.0));
        atomicAdd(&S[4 * NN + nd[e]], (unsigned long long)fxd(f * energy, 4294967296.0));
      }
    }
    if (a.node_q) {
      const double cc = charge;
      const double w[8] = {cc * (1 - ry) * (1 - rx) * (1 - rz), cc * (1 - ry) * rx * (1 - rz),
                           cc * (1 - ry) * (1 - rx) * rz,       cc * (1 - ry) * rx * rz,
                           cc * ry * (1 - rx) * (1 - rz),       cc * ry * rx * (1 - rz),
                           cc * ry * (1 - rx) * rz,             cc * ry * rx * rz};
#pragma unroll
      for (int e = 0; e < 8; ++e) atomicAdd(&a.node_q[nd[e]], (unsigned long long)fxd(w[e], 281474976710656.0));
    }
    if (a.cell_q) {
      atomicAdd(&a.cell_q[(pt == 0 ? 0 : NC) + c], (unsigned long long)fxd(charge, 281474976710656.0));
      atomicAdd(&a.cell_q[2 * NC + c], (unsigned long long)fxd(fabs(charge), 281474976710656.0));
    }
  }
}

// ------------------------------------------------------------------ RHS / fields
// particle_to_density normalisation (src/mcmodel.cpp:407-416) + compute_rhs (:2858-2893),
// eliminated form: rhs_u - sum over contact neighbours a_uv phi_v.
__global__ void k_rhs(RhsArgs a) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.nnode; n += stride) {
    double d = (double)(long long)a.node_q[n] / 281474976710656.0;
    const double v = a.pvol[n];
    if (fabs(v) > 0) d = d / v;
    double r = 0.0;
    if (fabs(d) > 0) r = d * v + a.pdop[n];
    const uint8_t kind = a.kind[n];
    if (kind == 1) r = a.dirichlet[n];
    a.rhs[n] = kind == 0 ? r - a.corr[n] : r;
    if (a.density) a.density[n] = d;
  }
}

// pot2field (src/mcmodel.cpp:2776-2813), Rho clamp (:2704-2711) and
// GetSurfRoughnessPhononScRate (src/GetSurfRoughnessPhononScRate.cpp:4-71), per cell.
__global__ void k_field(FieldArgs a) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const int cnx = a.cnx, cnz = a.cnz, pnx = a.pnx, pnz = a.pnz;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < a.ncell; c += stride) {
    const int k = (int)(c % cnz);
    const size_t t = c / cnz;
    const int i = (int)(t % cnx), j = (int)(t / cnx);
    const double* P = a.pot;
    auto p = [&](int ii, int jj, int kk) { return P[((size_t)jj * pnx + ii) * pnz + kk]; };
    const double p000 = p(i, j, k), p100 = p(i + 1, j, k), p010 = p(i, j + 1, k), p110 = p(i + 1, j + 1, k);
    const double p001 = p(i, j, k + 1), p101 = p(i + 1, j, k + 1), p011 = p(i, j + 1, k + 1),
                 p111 = p(i + 1, j + 1, k + 1);
    const double ex = -(p100 - p000 + p110 - p010 + p101 - p001 + p111 - p011) / (4.0 * a.dx[i]);
    const double ey = -(p010 + p110 + p011 + p111 - p000 - p100 - p001 - p101) / (4.0 * a.dy[j]);
    const double ez = -(p001 + p101 + p011 + p111 - p000 - p100 - p010 - p110) / (4.0 * a.dz[k]);
    a.ex[c] = ex; a.ey[c] = ey; a.ez[c] = ez;
    if (a.cell_q) {
      double v = (double)(long long)a.cell_q[2 * a.ncell + c] / 281474976710656.0;
      v /= a.cvol[c];
      const double lo = 0.5 * a.da[c];
      a.rho[c] = v > lo ? v : lo;
    }
    if (!a.rates) continue;
    const double Rho = a.rho[c];
    const int dir = a.eeff_dir[c];
    double Eeff = fabs(dir == 0 ? ex : dir == 1 ? ey : ez);
    if (Eeff < 1e-20) Eeff = 1e-20;
    for (int ipt = 0; ipt < 2; ++ipt) {
      const double effmass = a.mass_eff[ipt];
      double rr = 0.0, sp = 0.0;
      if (a.flag_rough)
        rr = 2 * PI * effmass * pow(a.surf_delta * a.surf_ail * Eeff * a.eps_si / (a.eps_si + a.eps_ox), 2);
      if (a.flag_sphonon) {
        const double Wcl = pow(a.Tn, a.surf_Nbmod) * 8.617343e-5 * a.T0 / (Eeff * a.field0);
        const double Wq = pow(12.0 * effmass * a.em0 * (Eeff * a.field0) / a.ec0 / a.hq0 / a.hq0, -1.0 / 3.0);
        const double Pf = a.surf_Pft * a.Tn + a.surf_Pfn * pow(Rho * a.conc0 / (1e17 * 1e6), a.surf_Npf) / a.Tn;
        const double Wclq = (a.surf_gama * Wcl + a.surf_theta * Wq) / Pf;
        sp = effmass * a.em0 * 8.617343e-5 * a.T0 * a.surf_XIph * a.surf_XIph /
                 (pow(a.hq0, 3) * a.sirho * pow(9050.02, 2) * Wclq) / a.scrt0 +
             a.ec0 / (effmass * a.em0 * a.surf_Kpha * pow(a.Tn, 2.24)) / a.scrt0;
      }
      a.rates[(2 * ipt) * a.ncell + c] = rr;
      a.rates[(2 * ipt + 1) * a.ncell + c] = sp;
    }
  }
}

// ------------------------------------------------------------------ PCG
// Operator: y = diag x + sum_dir c_dir(n) x(n+dir) + c_dir(n-dir) x(n-dir).
__device__ inline double apply_op(const PcgArgs& a, const double* __restrict__ x, size_t n, int i, int j, int k) {
  const size_t sx = a.pnz, sy = (size_t)a.pnx * a.pnz;
  double v = a.diag[n] * x[n];
  if (i + 1 < a.pnx) v += a.cx[n] * x[n + sx];
  if (i > 0) v += a.cx[n - sx] * x[n - sx];
  if (j + 1 < a.pny) v += a.cy[n] * x[n + sy];
  if (j > 0) v += a.cy[n - sy] * x[n - sy];
  if (k + 1 < a.pnz) v += a.cz[n] * x[n + 1];
  if (k > 0) v += a.cz[n - 1] * x[n - 1];
  return v;
}

// Block partial sums of three values into part[3*block + 0..2] (fixed order).
__device__ inline void block_reduce3(double s0, double s1, double s2, double* part) {
  __shared__ double sh[3][4];
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = s0; sh[1][w] = s1; sh[2][w] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { a0 += sh[0][q]; a1 += sh[1][q]; a2 += sh[2][q]; }
    part[3 * blockIdx.x] = a0;
    part[3 * blockIdx.x + 1] = a1;
    part[3 * blockIdx.x + 2] = a2;
  }
}

// Totals of the per-block partials in a fixed order; every block computes them.
__device__ inline void total3(const double* part, int nb, double* out) {
  __shared__ double sh[3][4];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) { a0 += part[3 * q]; a1 += part[3 * q + 1]; a2 += part[3 * q + 2]; }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = a0; sh[1][w] = a1; sh[2][w] = a2; }
  __syncthreads();
  out[0] = out[1] = out[2] = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { out[0] += sh[0][q]; out[1] += sh[1][q]; out[2] += sh[2][q]; }
  __syncthreads();
}

__device__ inline void node_ijk(const PcgArgs& a, size_t n, int& i, int& j, int& k) {
  k = (int)(n % a.pnz);
  const size_t t = n / a.pnz;
  i = (int)(t % a.pnx);
  j = (int)(t / a.pnx);
}

// fixed nodes take their value from b
__global__ __launch_bounds__(256) void k_pcg_fix(PcgArgs a) {
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.n; n += (size_t)gridDim.x * blockDim.x)
    if (a.kind[n] != 0) a.x[n] = a.b[n];
}

// r = b - A x, p = z = r/diag ; partials (r.z, r.r, b.b) -> pa
__global__ __launch_bounds__(256) void k_pcg_resid(PcgArgs a, double* pa) {
  double rz = 0, rr = 0, bb = 0;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.n; n += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    node_ijk(a, n, i, j, k);
    const double r = a.b[n] - apply_op(a, a.x, n, i, j, k);
    const double z = r / a.diag[n];
    a.r[n] = r;
    a.p[n] = z;
    rz += r * z;
    rr += r * r;
    bb += a.b[n] * a.b[n];
  }
  block_reduce3(rz, rr, bb, pa);
}

// scal[0] = r.r, scal[1] = r.z, scal[2] = b.b of the starting residual
__global__ __launch_bounds__(256) void k_pcg_norms(PcgArgs a, const double* pa) {
  double t[3];
  total3(pa, a.nblocks, t);
  if (threadIdx.x == 0) { a.scal[0] = t[1]; a.scal[1] = t[0]; a.scal[2] = t[2]; }
}

// q = A p ; partial p.q -> pb
__global__ __launch_bounds__(256) void k_pcg_spmv(PcgArgs a, double* pb) {
  double pq = 0;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.n; n += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    node_ijk(a, n, i, j, k);
    const double q = apply_op(a, a.p, n, i, j, k);
    a.q[n] = q;
    pq += a.p[n] * q;
  }
  block_reduce3(pq, 0.0, 0.0, pb);
}

// alpha = r.z / p.q ; x += alpha p ; r -= alpha q ; partials (r.z, r.r) -> pc
__global__ __launch_bounds__(256) void k_pcg_update(PcgArgs a, const double* pa, const double* pb, double* pc) {
  double ta[3], tb[3];
  total3(pa, a.nblocks, ta);
  total3(pb, a.nblocks, tb);
  const double alpha = (tb[0] != 0.0) ? ta[0] / tb[0] : 0.0;
  double rz = 0, rr = 0;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.n; n += (size_t)gridDim.x * blockDim.x) {
    a.x[n] += alpha * a.p[n];
    const double r = a.r[n] - alpha * a.q[n];
    a.r[n] = r;
    rz += r * (r / a.diag[n]);
    rr += r * r;
  }
  block_reduce3(rz, rr, 0.0, pc);
}

// beta = r.z(new) / r.z(old) ; p = z + beta p ; scal[0] = r.r, scal[1] = r.z
__global__ __launch_bounds__(256) void k_pcg_pupdate(PcgArgs a, const double* pa, const double* pc) {
  double ta[3], tc[3];
  total3(pa, a.nblocks, ta);
  total3(pc, a.nblocks, tc);
  const double beta = (ta[0] != 0.0) ? tc[0] / ta[0] : 0.0;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.n; n += (size_t)gridDim.x * blockDim.x)
    a.p[n] = a.r[n] / a.diag[n] + beta * a.p[n];
  if (blockIdx.x == 0 && threadIdx.x == 0) { a.scal[0] = tc[1]; a.scal[1] = tc[0]; }
}

// ------------------------------------------------------------------ helpers
__global__ void k_zero_u64(unsigned long long* p, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = 0;
}

}  // namespace emc

// ---------------------------------------------------------------- launchers
using namespace emc;

static int grid_for(long long n, int bs, int maxg = 65536) {
  long long g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > maxg) g = maxg;
  return (int)g;
}

// a.nslow must read zero on entry; the slow grid covers every particle, blocks
// past the listed count exit at once
hipError_t emc_launch_fly(const FlyArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fly_fast, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_fly_slow, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_fly_clones(const FlyArgs& a, unsigned long long b, unsigned long long e, hipStream_t s) {
  if (e <= b) return hipSuccess;
  hipLaunchKernelGGL(k_fly_clones, dim3(grid_for((long long)(e - b), 256)), dim3(256), 0, s, a, b, e);
  return hipGetLastError();
}
hipError_t emc_launch_merge(const FlyArgs& a, unsigned long long nclones, unsigned long long ndead,
                            unsigned long long* ticket, hipStream_t s) {
  if (nclones == 0) return hipSuccess;
  hipLaunchKernelGGL(k_merge, dim3(grid_for((long long)nclones, 256)), dim3(256), 0, s, a, nclones, ndead, ticket);
  return hipGetLastError();
}
hipError_t emc_launch_deposit(const DepositArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_deposit, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_rhs(const RhsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_rhs, dim3(grid_for((long long)a.nnode, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_field(const FieldArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_field, dim3(grid_for((long long)a.ncell, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_zero(unsigned long long* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for((long long)n, 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}
int emc_pcg_blocks(size_t n) { return grid_for((long long)n, 256, 1024); }
hipError_t emc_launch_pcg_start(const PcgArgs& a, hipStream_t s) {
  const int g = a.nblocks;
  hipLaunchKernelGGL(k_pcg_fix, dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_pcg_resid, dim3(g), dim3(256), 0, s, a, a.part);
  hipLaunchKernelGGL(k_pcg_norms, dim3(1), dim3(256), 0, s, a, (const double*)a.part);
  return hipGetLastError();
}
// iteration `it` ping-pongs the r.z partials between part (even) and part3 (odd)
hipError_t emc_launch_pcg_iter(const PcgArgs& a, int it, hipStream_t s) {
  const int g = a.nblocks;
  double* pa = (it & 1) ? a.part3 : a.part;
  double* pc = (it & 1) ? a.part : a.part3;
  hipLaunchKernelGGL(k_pcg_spmv, dim3(g), dim3(256), 0, s, a, a.part2);
  hipLaunchKernelGGL(k_pcg_update, dim3(g), dim3(256), 0, s, a, (const double*)pa, (const double*)a.part2, pc);
  hipLaunchKernelGGL(k_pcg_pupdate, dim3(g), dim3(256), 0, s, a, (const double*)pa, (const double*)pc);
  return hipGetLastError();
}
