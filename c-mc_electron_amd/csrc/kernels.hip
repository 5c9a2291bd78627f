// HIP kernels of the MI355X ensemble Monte Carlo step (gfx950, wave64).
//
//   k_fly_fast   one lane per particle: those whose step has only crossings
//                finish here, the rest are listed for
//   k_fly_slow   one lane per listed particle, full-band event loop (fly_core.h)
//   k_fly_clones same for GENERATE clones of this step
//   k_merge      surviving clones -> holes left by caught particles, then append
//   k_tile_*     counting sort of the particles by 4x8x4-cell tile (drops empty slots)
//   k_deposit_*  CIC charge, |q|-weighted node statistics, NGP cell sums: LDS per
//                tile, then a per-node gather of the tile blocks (int64, no global atomics)
//   k_rhs        node density and Poisson right-hand side (eliminated system)
//   k_field      cell fields from node potential, Rho clamp, surface rates
//   k_pcg_*      CG on the 7-point operator, preconditioned by
//   k_mg_*       one multigrid cycle: V(2,2) on the finest level, V(1,1) below
//                (damped Jacobi, Galerkin coarse levels, dense coarsest solve)
//
// All HBM-bound integer/FP64 work: no MFMA.  Compiled with -ffp-contract=off so
// the FP64 results match the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fly_core.h"
#include "kernels.h"
#include "qc.h"

namespace emc {

// ------------------------------------------------------------------ argument blocks
// A kernel takes its argument block by value (one device) or a device table of
// them, one per device of a batch that shares the mesh: one launch then covers
// every device, the device being blockIdx.y, and each block runs exactly the
// code of the single-device launch on its device's block.
template <class A>
__device__ __forceinline__ const A& kargs(const A& a) { return a; }
template <class A>
__device__ __forceinline__ const A& kargs(const A* t) { return t[blockIdx.y]; }

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
  f.left_time = f.M.dt;
  f.nclone = 0; f.flag_heat = a.flag_heat;
  f.n_events = 0; f.n_phonon = 0; f.n_impurity = 0;
  f.rng.init(f.key, a.step);
}

// Fast pass: particles whose step has only tetrahedron and plain cell crossings
// (Flight::fast_step) finish here.  The others are listed, wave-compacted, for
// k_fly_slow, which restarts them from the untouched state: the RNG is counter
// based, so the full event loop draws the same numbers and the result equals one
// pass of fly().  (Refilling finished lanes from a per-wave chunk, so that a
// wave does not wait for its longest history, was measured slower: most
// particles end after one or two events, so nearly every iteration paid for a
// refill.)
template <class AS>
__global__ __launch_bounds__(256, 3) void k_fly_fast(AS as_) {
  const FlyArgs& a = kargs(as_);
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
        Flight f(a.C->B, a.C->M, a.C->F, a.T);
        load_particle(f, a, q, pk);
        Flight::Loop L;
        f.fly_begin(L);
        int pend_kind = 0, pend_j = 0, r = Flight::FAST_DONE;
        while (f.left_time > 0) {
          r = f.fast_step(L, pend_kind, pend_j);
          if (r != Flight::FAST_MORE) break;
        }
        if (r != Flight::FAST_BAIL) {
          ev += f.n_events;
          f.commit_pending(pend_kind, pend_j);
          a.P.x[q] = f.x; a.P.y[q] = f.y; a.P.z[q] = f.z;
          a.P.kx[q] = f.kx; a.P.ky[q] = f.ky; a.P.kz[q] = f.kz;
          a.P.energy[q] = f.energy;
          const uint32_t pk2 = pk_make(f.par_type, f.icell, f.jcell, f.kcell);
          if (pk2 != pk) a.P.pk[q] = pk2;
          a.P.itet[q] = f.itet; a.P.isym[q] = f.isym;
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

// get_state() for the lanes of a wave that want a final state, with the wave's
// 64 lanes sharing a search that runs long.  Each lane first makes up to
// GS_SERIAL tries of its own (the common case ends there).  The lanes still
// drawing are then served one at a time: the lane's search is broadcast, every
// lane forms two positions of a chunk of 128 of its draws (lane j: Philox block
// j of the chunk, its words spread so that lane j holds positions j and j + 64)
// and evaluates them, and Flight::gs_walk replays the loop over the ballots.
// The draws, tries, tetrahedron and tallies are get_state()'s
// (Flight::get_state_chunked restates this on one thread for the host check).
// All lanes of the wave must call it together.
constexpr int GS_SERIAL = 1;
__device__ __forceinline__ void get_state_wave(Flight& f, bool want) {
  if (!__ballot(want)) return;
  const int lane = threadIdx.x & 63;
  double maxdostet = 0;
  int nted = 0, base = 0;
  bool small = false;
  const bool drew = want && f.gs_begin(maxdostet, nted, base, small);
  bool again = drew;
  int count2 = 0;
  while (__ballot(again && count2 < GS_SERIAL)) {
    if (!(again && count2 < GS_SERIAL)) continue;
    again = f.gs_try(f.energy, maxdostet, nted, base, small, count2);
  }
  for (unsigned long long need = __ballot(again); need; need &= need - 1) {
    const int L = __builtin_ctzll(need);
    const double e = __shfl(f.energy, L, 64), umax = __shfl(maxdostet, L, 64);
    const int unted = __shfl(nted, L, 64), ubase = __shfl(base, L, 64);
    const bool usmall = __shfl((int)small, L, 64) != 0;
    const uint32_t k0 = __shfl(f.rng.k0, L, 64), k1 = __shfl(f.rng.k1, L, 64);
    const uint32_t c1 = __shfl(f.rng.c1, L, 64), c2 = __shfl(f.rng.c2, L, 64);
    uint32_t blk = __shfl(f.rng.block, L, 64);
    int have = __shfl(f.rng.have, L, 64);
    double spare = __shfl(f.rng.spare, L, 64);
    int c = __shfl(count2, L, 64), fin = 0, nneg = 0;
    bool done = false;
    while (!done) {
      uint32_t w[4] = {blk + (uint32_t)lane, c1, c2, 0u};
      philox4x32_10(w, k0, k1);
      const double wa = Flight::gs_word(w, 0), wb = Flight::gs_word(w, 1);
      // position q: draw q - have of the blocks (word (q - have) & 1 of block
      // (q - have) / 2, held by that lane), or the spare draw
      double u[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = lane + 64 * h - have;
        const int src = r < 0 ? 0 : r >> 1;
        const double va = __shfl(wa, src, 64), vb = __shfl(wb, src, 64);
        u[h] = r < 0 ? spare : ((r & 1) ? vb : va);
      }
      int it[2];
      bool sp[2], ng[2];
      double dost[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) f.gs_eval(e, unted, ubase, usmall, u[h], it[h], sp[h], ng[h], dost[h]);
      // the draw after position q: lane + 1's, across the halves at lane 63
      const double n0a = __shfl(u[0], (lane + 1) & 63, 64), n0b = __shfl(u[1], 0, 64);
      const double n1 = __shfl(u[1], (lane + 1) & 63, 64);
      const double nx0 = lane < 63 ? n0a : n0b;
      const bool acc0 = sp[0] && !((nx0 * umax) > dost[0]);
      const bool acc1 = sp[1] && lane < 63 && !((n1 * umax) > dost[1]);
      const Flight::GsWalk wk = Flight::gs_walk(__ballot(sp[0]), __ballot(sp[1]), __ballot(acc0), __ballot(acc1),
                                                __ballot(ng[0]), __ballot(ng[1]), c);
      nneg += wk.nneg;
      c = wk.count2;
      if (wk.done) {
        fin = __shfl(wk.pos < 64 ? it[0] : it[1], wk.pos & 63, 64);
        done = true;
      }
      const int k = wk.consumed;
      const bool keep = have && k == 0;
      const int oldh = have;
      Flight::gs_advance(blk, have, k);
      if (have && !keep) spare = __shfl(wb, (k - 1 - oldh) / 2, 64);
    }
    if (lane == L) {
      f.itet = fin;
      count2 = c;
      f.rng.block = blk;
      f.rng.have = have;
      f.rng.spare = spare;
      again = false;
      if (nneg && f.T.counters) tally_add(&f.T.counters[CNT_DROP0 + 6], nneg);
    }
  }
  if (drew) f.gs_end(count2);
}

// Flight::fly() with the event loop run until the wave's last lane leaves it, so
// that get_state_wave has every lane.  Lanes with act false take part without a
// particle.
__device__ __forceinline__ void fly_wave(Flight& f, bool act) {
  Flight::Loop L;
  if (act) f.fly_begin(L);
  bool alive = act && f.left_time > 0;
  while (__ballot(alive)) {
    int flag = 0, post = Flight::POST_NONE;
    Flight::Saved s0{};
    if (alive && !f.event_head(L, flag, post, s0)) alive = false;
    get_state_wave(f, alive && post != Flight::POST_NONE);
    if (alive) alive = f.event_tail(L, flag, post, s0) && f.left_time > 0;
  }
}

// Full event loop for the particles k_fly_fast listed.  These histories diverge
// (scattering, zone faces, contacts), and a wave runs its lanes' different paths
// one after the other, so the list is spread thinly: each wave takes `lpw` entries
// (lanes 0..lpw-1), with lpw chosen so that the list covers about a.wave_slots
// waves, the number the GPU runs at once.
template <class AS>
__global__ __launch_bounds__(256) void k_fly_slow(AS as_) {
  const FlyArgs& a = kargs(as_);
  long long ev = 0, ph = 0, im = 0, ca = 0;
  const long long ns = (long long)*a.nslow;
  const long long slots = a.wave_slots > 0 ? a.wave_slots : 1;
  const long long lpw = ns <= slots ? 1 : (ns + slots - 1) / slots < 64 ? (ns + slots - 1) / slots : 64;
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  // every lane of the wave runs the loop (fly_wave needs them all); lanes from
  // lpw on, and lanes past the list, fly nothing
  for (long long e = wave * lpw + lane;; e += nwaves * lpw) {
    const bool act = lane < lpw && e < ns;
    if (!__ballot(act)) break;
    long long q = 0;
    Flight f(a.C->B, a.C->M, a.C->F, a.T);
    if (act) {
      q = a.slow[e];
      load_particle(f, a, q, a.P.pk[q]);
    }
    fly_wave(f, act);
    if (!act) continue;
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

// Clones [begin, end) of the clone buffer fly the remaining time of the step;
// `lpw` clones per wave, as in k_fly_slow.
__device__ __forceinline__ void fly_clone_range(const FlyArgs& a, unsigned long long begin, unsigned long long end, int lpw) {
  long long ev = 0, ph = 0, im = 0, ca = 0;
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  if (lane < lpw) {
    for (long long q = (long long)begin + wave * lpw + lane; q < (long long)end; q += nwaves * lpw) {
      CloneRec& c = a.clones[q];
      Flight f(a.C->B, a.C->M, a.C->F, a.T);
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
  }
  flush_counters(a.T, ev, ph, im, ca);
}

// clones [begin, end): the host-driven pass
__global__ __launch_bounds__(256) void k_fly_clones(FlyArgs a, unsigned long long begin, unsigned long long end,
                                                    int lpw) {
  fly_clone_range(a, begin, end, lpw);
}

// Clone passes without the host: ctr[4] .. ctr[5] is the range the next pass
// flies.  k_clone_range moves it on to the clones created since and counts the
// non-empty ranges in ctr[6] (one thread).
template <class AS>
__global__ void k_clone_range(AS as_) {
  const FlyArgs& a = kargs(as_);
  unsigned long long* ctr = a.nclones;
  const unsigned long long cap = a.clone_cap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr[4] = ctr[5];
    ctr[5] = ctr[0] < cap ? ctr[0] : cap;
    if (ctr[5] > ctr[4]) ctr[6] += 1;
  }
}

// k_fly_clones over ctr[4] .. ctr[5], spread as k_fly_slow spreads its list
template <class AS>
__global__ __launch_bounds__(256) void k_fly_clones_dev(AS as_) {
  const FlyArgs& a = kargs(as_);
  const unsigned long long* ctr = a.nclones;
  const long long begin = (long long)ctr[4], n = (long long)ctr[5] - begin;
  if (n <= 0) return;
  const long long slots = a.wave_slots > 0 ? a.wave_slots : 1;
  const int lpw = (int)(n <= slots ? 1 : ((n + slots - 1) / slots < 64 ? (n + slots - 1) / slots : 64));
  fly_clone_range(a, (unsigned long long)begin, (unsigned long long)(begin + n), lpw);
}

// Surviving clones take the slots of caught particles first, then go past n:
// survivors of clones [begin, end), ticket counts the survivors placed so far
// (ndead: *ndead_p)
__global__ __launch_bounds__(256) void k_merge(FlyArgs a, unsigned long long begin, unsigned long long end,
                                               const unsigned long long* ndead_p, unsigned long long* ticket) {
  const unsigned long long ndead = *ndead_p;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)begin + (long long)blockIdx.x * blockDim.x + threadIdx.x; q < (long long)end; q += stride) {
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

// k_merge over the clones flown by the device-driven passes: [0, ctr[4])
template <class AS>
__global__ __launch_bounds__(256) void k_merge_dev(AS as_) {
  const FlyArgs& a = kargs(as_);
  unsigned long long* ctr = a.nclones;
  const unsigned long long end = ctr[4], ndead = ctr[1];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < (long long)end; q += stride) {
    const CloneRec& c = a.clones[q];
    if (c.pk == PK_DEAD) continue;
    const unsigned long long m = atomicAdd(&ctr[2], 1ull);
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

// ------------------------------------------------------------------ tile sort
__device__ inline int tile_of(const SortArgs& a, uint32_t pk) {
  return ((pk_j(pk) / TILE_J) * a.ntx + pk_i(pk) / TILE_I) * a.ntz + pk_k(pk) / TILE_K;
}

// Lanes holding the same key in consecutive positions form a run; the run's first
// lane adds the run length to ctr[key] with one atomic and every lane of the run
// gets old value + its rank in the run.  key < 0 takes no part.  All 64 lanes of
// the wave must call it together.
// The same counting without the positions: no lane waits for the atomics.
__device__ inline void wave_run_count(unsigned* ctr, int key) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const unsigned long long above = heads & ~upto;
  const int nh = above ? __builtin_ctzll(above) : 64;
  if (head && key >= 0) (void)atomicAdd(&ctr[key], (unsigned)(nh - lane));
}

__device__ inline unsigned wave_run_add(unsigned* ctr, int key) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int hl = 63 - __builtin_clzll(heads & upto);
  const unsigned long long above = heads & ~upto;
  const int nh = above ? __builtin_ctzll(above) : 64;
  unsigned base = 0;
  if (head && key >= 0) base = atomicAdd(&ctr[key], (unsigned)(nh - lane));
  base = __shfl(base, hl, 64);
  return base + (unsigned)(lane - hl);
}

// Blocks 0 .. ntile-1 take the segment of their tile; the blocks after them
// stride over the slots past the segments (every particle there is a mover).
__device__ inline void seg_bounds(const SortArgs& a, const unsigned* off, unsigned& lo, unsigned& hi) {
  const int t = blockIdx.x;
  lo = hi = 0;
  if (a.seg_valid) {
    const unsigned n = (unsigned)a.n;
    lo = min(off[t], n);
    hi = min(off[t + 1], n);
  }
}
__device__ inline long long append_begin(const SortArgs& a, const unsigned* off) {
  return a.seg_valid ? (long long)off[a.ntile] : 0;
}

// Stayers of each segment (one plain store per tile) and movers per destination
// tile (an atomic per run of equal tiles in a wave).
template <class AS>
__global__ __launch_bounds__(256) void k_tile_hist(AS as_) {
  const SortArgs& a = kargs(as_);
  const int lane = threadIdx.x & 63;
  if ((int)blockIdx.x < a.ntile) {
    __shared__ unsigned stay_sum;
    if (threadIdx.x == 0) stay_sum = 0;
    __syncthreads();
    unsigned lo, hi;
    seg_bounds(a, a.toff, lo, hi);
    unsigned stay = 0;  // lane 0 of each wave
    for (unsigned base = lo; base < hi; base += blockDim.x) {
      const unsigned q = base + threadIdx.x;
      int t = -1;
      bool st = false;
      if (q < hi) {
        const uint32_t pk = a.P.pk[q];
        if (pk != PK_DEAD) {
          t = tile_of(a, pk);
          st = t == (int)blockIdx.x;
        }
      }
      const unsigned long long m = __ballot(st);
      if (lane == 0) stay += (unsigned)__popcll(m);
      wave_run_count(a.tcount, st ? -1 : t);
    }
    if (lane == 0 && stay) atomicAdd(&stay_sum, stay);
    __syncthreads();
    if (threadIdx.x == 0) a.tstay[blockIdx.x] = stay_sum;
    return;
  }
  const long long stride = (long long)(gridDim.x - a.ntile) * blockDim.x;
  for (long long base = append_begin(a, a.toff) + (long long)(blockIdx.x - a.ntile) * blockDim.x; base < a.n;
       base += stride) {
    const long long q = base + threadIdx.x;
    int t = -1;
    if (q < a.n) {
      const uint32_t pk = a.P.pk[q];
      if (pk != PK_DEAD) t = tile_of(a, pk);
    }
    wave_run_count(a.tcount, t);
  }
}

// One block of 1024 threads: the tile counts (stayers + movers) are scanned into
// toff, whose previous contents (the segments the scatter reads) go to toff_old;
// the mover counts and cursors are cleared for the next sort.
template <class AS>
__global__ __launch_bounds__(1024) void k_tile_scan(AS as_) {
  const SortArgs& a = kargs(as_);
  __shared__ unsigned wsum[16];
  const int nt = a.ntile;
  const int tid = threadIdx.x;
  const int chunk = (nt + 1023) / 1024;
  const int b0 = tid * chunk, b1 = min(nt, b0 + chunk);
  const unsigned old_total = a.toff[nt];
  unsigned loc = 0;
  for (int t = b0; t < b1; ++t) loc += a.tstay[t] + a.tcount[t];
  // inclusive scan of loc over the block
  const int lane = tid & 63, w = tid >> 6;
  unsigned v = loc;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  unsigned woff = 0;
  for (int q = 0; q < w; ++q) woff += wsum[q];
  unsigned run = woff + v - loc;  // exclusive prefix of this thread's chunk
  for (int t = b0; t < b1; ++t) {
    const unsigned cnt = a.tstay[t] + a.tcount[t];
    a.toff_old[t] = a.seg_valid ? a.toff[t] : 0u;
    a.toff[t] = run;
    run += cnt;
    a.tcount[t] = 0;
    a.tcursor[t] = 0;
  }
  if (tid == 1023) {
    unsigned tot = 0;
    for (int q = 0; q < 16; ++q) tot += wsum[q];
    a.toff_old[nt] = a.seg_valid ? old_total : 0u;
    a.toff[nt] = tot;
  }
}

// Stayers of tile t land at toff[t] + their rank among the segment's stayers
// (a block scan: no atomics, the segment's order kept); movers into tile t at
// toff[t] + tstay[t] + a cursor (one atomic per run of equal tiles in a wave).
// Each particle is read before its slot is known, so the loads overlap the
// cursor atomic's round trip.  512 threads, so that most segments (a few hundred
// particles) take one pass of the loop.
template <class AS>
__global__ __launch_bounds__(512) void k_tile_scatter(AS as_) {
  const SortArgs& a = kargs(as_);
  const bool seg = (int)blockIdx.x < a.ntile;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __shared__ unsigned wtot[16];
  long long base, end, stride;
  unsigned ranked = 0;  // stayers of the segment placed so far
  unsigned dst0 = 0;
  if (seg) {
    unsigned lo, hi;
    seg_bounds(a, a.toff_old, lo, hi);
    base = lo;
    end = hi;
    stride = blockDim.x;
    dst0 = a.toff[blockIdx.x];
  } else {
    base = append_begin(a, a.toff_old) + (long long)(blockIdx.x - a.ntile) * blockDim.x;
    end = a.n;
    stride = (long long)(gridDim.x - a.ntile) * blockDim.x;
  }
  for (; base < end; base += stride) {
    const long long q = base + threadIdx.x;
    int t = -1;
    uint32_t pk = PK_DEAD;
    double x = 0, y = 0, z = 0, kx = 0, ky = 0, kz = 0, en = 0, ch = 0;
    int32_t itet = 0, isym = 0;
    uint64_t key = 0;
    if (q < end) {
      // every field loaded with pk (an empty slot's are read and dropped), so the
      // loads do not wait for pk
      pk = a.P.pk[q];
      x = a.P.x[q]; y = a.P.y[q]; z = a.P.z[q]; kx = a.P.kx[q]; ky = a.P.ky[q]; kz = a.P.kz[q];
      en = a.P.energy[q]; ch = a.P.charge[q]; itet = a.P.itet[q]; isym = a.P.isym[q]; key = a.P.key[q];
      if (pk != PK_DEAD) t = tile_of(a, pk);
    }
    const bool st = seg && t == (int)blockIdx.x;
    const unsigned r = wave_run_add(a.tcursor, st ? -1 : t);
    size_t d = 0;
    if (seg) {
      // rank of this stayer: stayers of the earlier waves, then of the lanes below
      const unsigned long long m = __ballot(st);
      if (lane == 0) wtot[w] = (unsigned)__popcll(m);
      __syncthreads();
      unsigned before = 0, total = 0;
      for (int o = 0; o < nw; ++o) {
        const unsigned c = wtot[o];
        if (o < w) before += c;
        total += c;
      }
      __syncthreads();
      if (st) d = (size_t)dst0 + ranked + before + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
      ranked += total;
    }
    if (t < 0) continue;
    if (!st) d = (size_t)a.toff[t] + a.tstay[t] + r;
    a.Q.x[d] = x; a.Q.y[d] = y; a.Q.z[d] = z;
    a.Q.kx[d] = kx; a.Q.ky[d] = ky; a.Q.kz[d] = kz;
    a.Q.energy[d] = en; a.Q.charge[d] = ch;
    a.Q.pk[d] = pk; a.Q.itet[d] = itet; a.Q.isym[d] = isym; a.Q.key[d] = key;
  }
}

// compute_rhs at node n from its charge sum q (src/mcmodel.cpp:2858-2893)
__device__ inline void rhs_node(const RhsArgs& a, size_t n, unsigned long long q) {
  double d = (double)(long long)q / 281474976710656.0;
  const double v = a.pvol[n];
  if (fabs(v) > 0) d = d / v;
  double r = 0.0;
  if (fabs(d) > 0) r = d * v + a.pdop[n];
  const uint8_t kind = a.kind[n];
  if (kind == 1) r = a.dirichlet[n];
  a.rhs[n] = kind == 0 ? r - a.corr[n] : r;
  if (a.density) a.density[n] = d;
}

// statistic()'s window sums at node q
__device__ inline void stat_node(const StatArgs& a, size_t q) {
  a.stat_pot[q] += a.pot[q];
  if (a.heat_e) {
    a.stat_heat_e[q] += a.heat_e[q];
    a.stat_heat_h[q] += a.heat_h[q];
    if (a.zero_heat) {
      a.heat_e[q] = 0;
      a.heat_h[q] = 0;
    }
  }
  if (a.qc_pot) a.stat_qc_pot[q] += a.qc_pot[q];
}

// ------------------------------------------------------------------ deposit
// statistic (src/mcmodel.cpp:1167-1408), particle_to_density (:348-417) and
// compute_cell_charge (:2664-2712) in one pass over the tile-sorted particles.
// One workgroup per tile sums the fixed-point contributions in LDS; each cell
// belongs to one tile and is stored directly, the node block goes to `blk` and
// k_deposit_gather adds the (up to 8) blocks that share a node.  Integer sums:
// the result does not depend on the order of the particles.  In a step the
// gather also forms the Poisson right-hand side and statistic()'s potential and
// heat sums of each node, and the cell pass adds the cell counts to the window.
template <class AS>
__global__ __launch_bounds__(256) void k_deposit_tile(AS as_) {
  const DepositArgs& a = kargs(as_);
  __shared__ unsigned long long sn[DEP_NQ * TILE_NODES];
  __shared__ unsigned long long sc[4 * TILE_CELLS];
  const int t = blockIdx.x;
  const int tk = t % a.ntz, ti = (t / a.ntz) % a.ntx, tj = t / (a.ntz * a.ntx);
  const int i0 = ti * TILE_I, j0 = tj * TILE_J, k0 = tk * TILE_K;
  const unsigned p0 = a.toff[t], p1 = a.toff[t + 1];
  const size_t NC = a.ncell;
  for (int e = threadIdx.x; e < DEP_NQ * TILE_NODES; e += blockDim.x) sn[e] = 0;
  for (int e = threadIdx.x; e < 4 * TILE_CELLS; e += blockDim.x) sc[e] = 0;
  __syncthreads();
  const bool want_stat = a.node_stat != nullptr;
  // Lane l of wave w takes entry 4l + w of each run of 256: the particles of one
  // cell tend to sit together in the tile, and spreading them over the waves
  // keeps the lanes of a wave off the same LDS words (same-address atomics
  // serialise).
  const unsigned spread = ((threadIdx.x & 63u) << 2) | (threadIdx.x >> 6);
  for (unsigned q0 = p0; q0 < p1; q0 += blockDim.x) {
    const unsigned q = q0 + (blockDim.x == 256 ? spread : threadIdx.x);
    if (q >= p1) continue;
    const uint32_t pk = a.P.pk[q];
    const int i = pk_i(pk), j = pk_j(pk), k = pk_k(pk), pt = pk_type(pk);
    const double x = a.P.x[q], y = a.P.y[q], z = a.P.z[q], charge = a.P.charge[q], energy = a.P.energy[q];
    const double rx = (x - a.M.lx[i]) / a.M.dx[i];
    const double ry = (y - a.M.ly[j]) / a.M.dy[j];
    const double rz = (z - a.M.lz[k]) / a.M.dz[k];
    const int li = i - i0, lj = j - j0, lk = k - k0;
    auto lnode = [&](int di, int dj, int dk) { return ((lj + dj) * (TILE_I + 1) + (li + di)) * (TILE_K + 1) + (lk + dk); };
    const int nd[8] = {lnode(0, 0, 0), lnode(1, 0, 0), lnode(0, 0, 1), lnode(1, 0, 1),
                       lnode(0, 1, 0), lnode(1, 1, 0), lnode(0, 1, 1), lnode(1, 1, 1)};
    const int lc = (lj * TILE_I + li) * TILE_K + lk;
    atomicAdd(&sc[3 * TILE_CELLS + lc], 1ull);
    if (want_stat) {
      const int isym = a.P.isym[q];
      const double* vt = a.B.tr[a.P.itet[q]].vt;
      const int* ms = a.B.matsym;
      const double v3[3] = {vt[0], vt[1], vt[2]};
      const double vx = v3[ms[0 * 48 + isym] - 1] * ms[3 * 48 + isym];
      const double vy = v3[ms[1 * 48 + isym] - 1] * ms[4 * 48 + isym];
      const double vz = v3[ms[2 * 48 + isym] - 1] * ms[5 * 48 + isym];
      unsigned long long* S = sn + (size_t)(pt == 0 ? 1 : 6) * TILE_NODES;
      const double cc[8] = {charge * (1 - rx) * (1 - ry) * (1 - rz), charge * rx * (1 - ry) * (1 - rz),
                            charge * (1 - rx) * (1 - ry) * rz,       charge * rx * (1 - ry) * rz,
                            charge * (1 - rx) * ry * (1 - rz),       charge * rx * ry * (1 - rz),
                            charge * (1 - rx) * ry * rz,             charge * rx * ry * rz};
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const double f = fabs(cc[e]);
        atomicAdd(&S[0 * TILE_NODES + nd[e]], (unsigned long long)fxd(vx * f, 4294967296.0));
        atomicAdd(&S[1 * TILE_NODES + nd[e]], (unsigned long long)fxd(vy * f, 4294967296.0));
        atomicAdd(&S[2 * TILE_NODES + nd[e]], (unsigned long long)fxd(vz * f, 4294967296.0));
        atomicAdd(&S[3 * TILE_NODES + nd[e]], (unsigned long long)fxd(f, 4294967296.0));
        atomicAdd(&S[4 * TILE_NODES + nd[e]], (unsigned long long)fxd(f * energy, 4294967296.0));
      }
    }
    {
      const double cc = charge;
      const double w[8] = {cc * (1 - ry) * (1 - rx) * (1 - rz), cc * (1 - ry) * rx * (1 - rz),
                           cc * (1 - ry) * (1 - rx) * rz,       cc * (1 - ry) * rx * rz,
                           cc * ry * (1 - rx) * (1 - rz),       cc * ry * rx * (1 - rz),
                           cc * ry * (1 - rx) * rz,             cc * ry * rx * rz};
#pragma unroll
      for (int e = 0; e < 8; ++e) atomicAdd(&sn[nd[e]], (unsigned long long)fxd(w[e], 281474976710656.0));
    }
    atomicAdd(&sc[(pt == 0 ? 0 : TILE_CELLS) + lc], (unsigned long long)fxd(charge, 281474976710656.0));
    atomicAdd(&sc[2 * TILE_CELLS + lc], (unsigned long long)fxd(fabs(charge), 281474976710656.0));
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TILE_CELLS; e += blockDim.x) {
    const int lk = e % TILE_K, li = (e / TILE_K) % TILE_I, lj = e / (TILE_K * TILE_I);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    if (i >= a.M.cnx || j >= a.M.cny || k >= a.M.cnz) continue;
    const size_t c = ((size_t)j * a.M.cnx + i) * a.M.cnz + k;
    if (a.cell_n) a.cell_n[c] = sc[3 * TILE_CELLS + e];
    if (a.st.stat_enum) a.st.stat_enum[c] += sc[3 * TILE_CELLS + e];
    if (a.cell_q) {
      a.cell_q[c] = sc[e];
      a.cell_q[NC + c] = sc[TILE_CELLS + e];
      a.cell_q[2 * NC + c] = sc[2 * TILE_CELLS + e];
    }
  }
  if (p1 > p0) {
    // node-major in memory ([node][quantity]): the gather reads a node's sums
    // from one 88-byte run
    unsigned long long* b = a.blk + (size_t)t * DEP_NQ * TILE_NODES;
    for (int e = threadIdx.x; e < DEP_NQ * TILE_NODES; e += blockDim.x) {
      const int ln = e / DEP_NQ, q = e - ln * DEP_NQ;
      b[e] = sn[q * TILE_NODES + ln];
    }
  }
}

// One thread per node: sum the node's entries of the non-empty tiles around it.
template <class AS>
__global__ __launch_bounds__(256) void k_deposit_gather(AS as_) {
  const DepositArgs& a = kargs(as_);
  const size_t NN = a.nnode;
  const int pnx = a.M.pnx, pnz = a.M.pnz;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < NN; n += (size_t)gridDim.x * blockDim.x) {
    const int K = (int)(n % pnz), I = (int)((n / pnz) % pnx), J = (int)(n / ((size_t)pnz * pnx));
    int tx[2], lx[2], nx = 0, ty[2], ly[2], ny = 0, tz[2], lz[2], nz = 0;
    if (I / TILE_I < a.ntx) { tx[nx] = I / TILE_I; lx[nx++] = I % TILE_I; }
    if (I % TILE_I == 0 && I > 0) { tx[nx] = I / TILE_I - 1; lx[nx++] = TILE_I; }
    if (J / TILE_J < a.nty) { ty[ny] = J / TILE_J; ly[ny++] = J % TILE_J; }
    if (J % TILE_J == 0 && J > 0) { ty[ny] = J / TILE_J - 1; ly[ny++] = TILE_J; }
    if (K / TILE_K < a.ntz) { tz[nz] = K / TILE_K; lz[nz++] = K % TILE_K; }
    if (K % TILE_K == 0 && K > 0) { tz[nz] = K / TILE_K - 1; lz[nz++] = TILE_K; }
    unsigned long long s[DEP_NQ];
#pragma unroll
    for (int q = 0; q < DEP_NQ; ++q) s[q] = 0;
    for (int b = 0; b < ny; ++b)
      for (int aa = 0; aa < nx; ++aa)
        for (int c = 0; c < nz; ++c) {
          const int t = (ty[b] * a.ntx + tx[aa]) * a.ntz + tz[c];
          if (a.toff[t + 1] == a.toff[t]) continue;
          const int ln = (ly[b] * (TILE_I + 1) + lx[aa]) * (TILE_K + 1) + lz[c];
          const unsigned long long* blk = a.blk + ((size_t)t * TILE_NODES + ln) * DEP_NQ;
#pragma unroll
          for (int q = 0; q < DEP_NQ; ++q) s[q] += blk[q];
        }
    if (a.node_q) a.node_q[n] = s[0];
    if (a.rhs.rhs) rhs_node(a.rhs, n, s[0]);
    if (a.st.stat_pot) stat_node(a.st, n);
    if (a.node_stat) {
#pragma unroll
      for (int q = 0; q < 10; ++q) {
        unsigned long long* dst = a.node_stat + (size_t)q * NN + n;
        *dst = (a.stat_accumulate ? *dst : 0ull) + s[1 + q];
      }
    }
  }
}

// ------------------------------------------------------------------ RHS / fields
// particle_to_density normalisation (src/mcmodel.cpp:407-416) + compute_rhs (:2858-2893),
// eliminated form: rhs_u - sum over contact neighbours a_uv phi_v.
// One lane per node of the quantum region; the window's sources are read in the
// reference's loop order (i, then j, then k) and every term is formed as
// ((V + Eb) * wz) * wy * wx, so the sum is the oracle's to the last bit.
// Neighbouring lanes are neighbouring nodes (k fastest) whose windows overlap
// almost entirely: the loads are served by L2.
__global__ __launch_bounds__(256) void k_qc_potential(QcArgs a) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < a.nq; q += stride) {
    const int p = a.qnodes[q];
    const int pk = p % a.pnz;
    const int t = p / a.pnz;
    const int pi = t % a.pnx, pj = t / a.pnx;
    const int i0 = a.lo[0][pi], i1 = a.hi[0][pi];
    const int j0 = a.lo[1][pj], j1 = a.hi[1][pj];
    const int k0 = a.lo[2][pk], k1 = a.hi[2][pk];
    const double* WX = a.w[0] + (size_t)pi * a.wlen[0] * 4;
    const double* WY = a.w[1] + (size_t)pj * a.wlen[1] * 4;
    const double* WZ = a.w[2] + (size_t)pk * a.wlen[2] * 4;
    double ret = 0.0;
    for (int i = i0; i <= i1; ++i)
      for (int j = j0; j <= j1; ++j) {
        const size_t row = ((size_t)j * a.pnx + i) * a.pnz;
        for (int k = k0; k <= k1; ++k) {
          const uint16_t f = a.flag[row + k];
          if (!(f & (QN_OXIDE | QN_SILICON))) continue;
          const double eb = ((f & QN_OXIDE) && !(f & QN_SILICON)) ? -a.eb : 0.0;
          const double v = a.pot[row + k] + eb;
          const int sx = ((f & QN_UP) ? 1 : 0) | ((f & QN_DOWN) ? 2 : 0);
          const int sy = ((f & QN_LEFT) ? 1 : 0) | ((f & QN_RIGHT) ? 2 : 0);
          const int sz = ((f & QN_FRONT) ? 1 : 0) | ((f & QN_BACK) ? 2 : 0);
          ret += v * WZ[(k - k0) * 4 + sz] * WY[(j - j0) * 4 + sy] * WX[(i - i0) * 4 + sx];
        }
      }
    a.qc_pot[p] = ret * 0.125;
  }
}


__global__ void k_rhs(RhsArgs a) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.nnode; n += stride) {
    rhs_node(a, n, a.node_q[n]);
  }
}

// pot2field (src/mcmodel.cpp:2776-2813), Rho clamp (:2704-2711) and
// GetSurfRoughnessPhononScRate (src/GetSurfRoughnessPhononScRate.cpp:4-71), per cell.
template <class AS>
__global__ void k_field(AS as_) {
  const FieldArgs& a = kargs(as_);
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
// Keeps the compiler from moving memory accesses across this point: the loads
// above are all issued before the arithmetic below waits for the first of them
// (otherwise the scheduler interleaves load, wait, multiply-add per neighbour).
__device__ inline void mem_fence_compiler() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

template <class PA>
__device__ inline double apply_op(const PA& a, const typename PA::real* __restrict__ x, size_t n, int i, int j,
                                  int k) {
  // every neighbour is loaded from a clamped address and a missing one adds 0.0,
  // so the loads carry no branches and issue together (the sum is the same, term
  // for term, as over the present neighbours only)
  const size_t sx = a.pnz, sy = (size_t)a.pnx * a.pnz;
  const bool ip = i + 1 < a.pnx, im = i > 0, jp = j + 1 < a.pny, jm = j > 0, kp = k + 1 < a.pnz, km = k > 0;
  const size_t nim = im ? n - sx : n, njm = jm ? n - sy : n, nkm = km ? n - 1 : n;
  const double x0 = x[n], x1 = x[ip ? n + sx : n], x2 = x[nim], x3 = x[jp ? n + sy : n], x4 = x[njm],
               x5 = x[kp ? n + 1 : n], x6 = x[nkm];
  const double c0 = a.diag[n], c1 = a.cx[n], c2 = a.cx[nim], c3 = a.cy[n], c4 = a.cy[njm], c5 = a.cz[n], c6 = a.cz[nkm];
  mem_fence_compiler();
  double v = c0 * x0;
  v += ip ? c1 * x1 : 0.0;
  v += im ? c2 * x2 : 0.0;
  v += jp ? c3 * x3 : 0.0;
  v += jm ? c4 * x4 : 0.0;
  v += kp ? c5 * x5 : 0.0;
  v += km ? c6 * x6 : 0.0;
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

// Block partial sum of one value into part[block] (the order of block_reduce3's
// first component).
__device__ inline void block_reduce1(double s0, double* part) {
  __shared__ double sh[4];
  s0 = wave_sum(s0);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = s0;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a0 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) a0 += sh[q];
    part[blockIdx.x] = a0;
  }
}

// Total of one-value partials (block_reduce1) in total3's order.
__device__ inline double total1(const double* part, int nb) {
  __shared__ double sh[4];
  double a0 = 0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) a0 += part[q];
  a0 = wave_sum(a0);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = a0;
  __syncthreads();
  double t = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sh[q];
  __syncthreads();
  return t;
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

// Totals of two one-value partial arrays (block_reduce1), each summed in total3's
// order, in one pass.
__device__ inline void total_first2(const double* pa, const double* pb, int nb, double& ta, double& tb) {
  __shared__ double sh[2][4];
  double a0 = 0, b0 = 0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) { a0 += pa[q]; b0 += pb[q]; }
  a0 = wave_sum(a0);
  b0 = wave_sum(b0);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = a0; sh[1][w] = b0; }
  __syncthreads();
  ta = tb = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { ta += sh[0][q]; tb += sh[1][q]; }
  __syncthreads();
}

template <class PA>
__device__ inline void node_ijk(const PA& a, size_t n, int& i, int& j, int& k) {
  k = (int)(n % a.pnz);
  const size_t t = n / a.pnz;
  i = (int)(t % a.pnx);
  j = (int)(t / a.pnx);
}

#define GRID_LOOP(n, N) for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < (N); n += (size_t)gridDim.x * blockDim.x)

// fixed nodes take their value from b; the solve starts not done
template <class AS>
__global__ __launch_bounds__(256) void k_pcg_fix(AS as_) {
  const auto& a = kargs(as_);
  using R = typename std::decay_t<decltype(a)>::real;
  GRID_LOOP(n, a.n) if (a.kind[n] != 0) a.x[n] = a.b[n];
  if (blockIdx.x == 0 && threadIdx.x == 0) { *a.done = 0; a.scal[3] = 0; }
}

// r = b - A x ; partials (r.r, b.b) -> part_a
template <class AS>
__global__ __launch_bounds__(256) void k_pcg_resid(AS as_) {
  const auto& a = kargs(as_);
  using R = typename std::decay_t<decltype(a)>::real;
  double rr = 0, bb = 0;
  GRID_LOOP(n, a.n) {
    int i, j, k;
    node_ijk(a, n, i, j, k);
    const double r = (double)a.b[n] - apply_op(a, a.x, n, i, j, k);
    a.r[n] = (R)r;
    if (a.r32) a.r32[n] = (float)r;
    rr += r * r;
    bb += (double)a.b[n] * (double)a.b[n];
  }
  block_reduce3(rr, bb, 0.0, a.part_a);
}

// scal = (r.r, r.z, b.b, 0, rtol^2 b.b); done when the start meets the tolerance
// (one block)
template <class AS>
__global__ __launch_bounds__(256) void k_pcg_begin(AS as_) {
  const auto& a = kargs(as_);
  using R = typename std::decay_t<decltype(a)>::real;
  double ta[3];
  total3(a.part_a, a.nblocks, ta);
  const double tz = total1(a.part_rz0, a.nblocks);
  if (threadIdx.x == 0) {
    a.scal[0] = ta[0]; a.scal[1] = tz; a.scal[2] = ta[1];
    a.scal[4] = a.rtol * a.rtol * ta[1];
    if (ta[0] <= a.scal[4]) *a.done = 1;
  }
}

// p = z + beta p_in (p = z in the first iteration), beta = r.z(new) / r.z(old),
// then q = A p and the partial p.q.  p at the six neighbours is recomputed from
// z and p_in with the same operations, so p_out is written once and p_in never
// changes during the pass.
template <class AS>
__global__ __launch_bounds__(256) void k_pcg_spmvp(AS as_, int it) {
  const auto& a = kargs(as_);
  using R = typename std::decay_t<decltype(a)>::real;
  // iteration `it`: rz_new holds r.z of the last cycle (part_rz0 for even it),
  // rz_old that of the cycle before; p alternates between p and p2
  const double* rz_old = (it & 1) ? a.part_rz0 : a.part_rz1;
  const double* rz_new = (it & 1) ? a.part_rz1 : a.part_rz0;
  R* p_out = (it & 1) ? a.p2 : a.p;
  const R* p_in = (it & 1) ? a.p : a.p2;
  const int first = it == 0 ? 1 : 0;
  if (*a.done) return;
  double beta = 0.0;
  if (!first) {
    double to, tn;
    total_first2(rz_old, rz_new, a.nblocks, to, tn);
    beta = (to != 0.0) ? tn / to : 0.0;
  }
  const size_t sx = a.pnz, sy = (size_t)a.pnx * a.pnz;
  double pq = 0;
  // z in the CG's precision, or the FP32 cycle's output.  Every neighbour is
  // loaded from a clamped address and a missing one adds 0.0, so the loads of a
  // node issue together (as in apply_op).
  auto pass = [&](const auto* __restrict__ z, auto first_c) {
    constexpr bool FIRST = decltype(first_c)::value;
    GRID_LOOP(n, a.n) {
      int i, j, k;
      node_ijk(a, n, i, j, k);
      const bool ip = i + 1 < a.pnx, im = i > 0, jp = j + 1 < a.pny, jm = j > 0, kp = k + 1 < a.pnz, km = k > 0;
      const size_t nb[7] = {n, ip ? n + sx : n, im ? n - sx : n, jp ? n + sy : n, jm ? n - sy : n,
                            kp ? n + 1 : n, km ? n - 1 : n};
      double zv[7], pv[7];
#pragma unroll
      for (int o = 0; o < 7; ++o) zv[o] = z[nb[o]];
      if (!FIRST) {
#pragma unroll
        for (int o = 0; o < 7; ++o) pv[o] = p_in[nb[o]];
      }
      const double c0 = a.diag[n], c1 = a.cx[n], c2 = a.cx[nb[2]], c3 = a.cy[n], c4 = a.cy[nb[4]], c5 = a.cz[n],
                   c6 = a.cz[nb[6]];
      mem_fence_compiler();
      double p[7];
#pragma unroll
      for (int o = 0; o < 7; ++o) p[o] = FIRST ? zv[o] : zv[o] + beta * pv[o];
      double q = c0 * p[0];
      q += ip ? c1 * p[1] : 0.0;
      q += im ? c2 * p[2] : 0.0;
      q += jp ? c3 * p[3] : 0.0;
      q += jm ? c4 * p[4] : 0.0;
      q += kp ? c5 * p[5] : 0.0;
      q += km ? c6 * p[6] : 0.0;
      p_out[n] = (R)p[0];
      a.q[n] = (R)q;
      pq += p[0] * q;
    }
  };
  using T1 = std::integral_constant<bool, true>;
  using T0 = std::integral_constant<bool, false>;
  if (a.z32) {
    if (first) pass(a.z32, T1{});
    else pass(a.z32, T0{});
  } else {
    if (first) pass((const R*)a.z, T1{});
    else pass((const R*)a.z, T0{});
  }
  block_reduce1(pq, a.part_pq);
}

// alpha = r.z / p.q ; x += alpha p ; r -= alpha q ; partial r.r
template <class AS>
__global__ __launch_bounds__(256) void k_pcg_update(AS as_, int it) {
  const auto& a = kargs(as_);
  using R = typename std::decay_t<decltype(a)>::real;
  // r.z of the last cycle, and the search direction k_pcg_spmvp wrote
  const double* rz_old = (it & 1) ? a.part_rz1 : a.part_rz0;
  const R* p = (it & 1) ? a.p2 : a.p;
  if (*a.done) return;
  double ta, tb;
  total_first2(rz_old, a.part_pq, a.nblocks, ta, tb);
  const double alpha = (tb != 0.0) ? ta / tb : 0.0;
  double rr = 0;
  GRID_LOOP(n, a.n) {
    a.x[n] = (R)((double)a.x[n] + alpha * (double)p[n]);
    const double r = (double)a.r[n] - alpha * (double)a.q[n];
    a.r[n] = (R)r;
    if (a.r32) a.r32[n] = (float)r;
    rr += r * r;
  }
  block_reduce1(rr, a.part_rr);
}

// ------------------------------------------------------------------ multigrid V-cycle
// The phases of the cycle are written once per node (grid-stride loops, one
// launch per phase).  A single workgroup walking the small levels was measured
// slower (latency bound) than one launch per phase.
struct GridLoop {
  __device__ static size_t start() { return (size_t)blockIdx.x * blockDim.x + threadIdx.x; }
  __device__ static size_t stride() { return (size_t)gridDim.x * blockDim.x; }
};

template <class LV>
__device__ inline void lvl_ijk(const LV& L, size_t n, int& i, int& j, int& k) {
  k = (int)(n % L.nz);
  const size_t t = n / L.nz;
  i = (int)(t % L.nx);
  j = (int)(t / L.nx);
}

// (A_l x)(n) with x given by an accessor; level 0 is the 7-point operator in FP64,
// coarser levels the symmetric half stencil.  Every neighbour is loaded from a
// clamped address and a missing one adds 0.0, so the loads carry no branches and
// issue together; the sum is the same, term for term, as over the present
// neighbours only.
template <bool FINE, class MA, class LV, class XA>
__device__ inline double mg_Ax(const MA& m, const LV& L, const XA& xa, size_t n, int i, int j, int k) {
  if (FINE) {
    const size_t sx = L.nz, sy = (size_t)L.nx * L.nz;
    const bool ip = i + 1 < L.nx, im = i > 0, jp = j + 1 < L.ny, jm = j > 0, kp = k + 1 < L.nz, km = k > 0;
    const size_t nip = ip ? n + sx : n, nim = im ? n - sx : n, njp = jp ? n + sy : n, njm = jm ? n - sy : n;
    const size_t nkp = kp ? n + 1 : n, nkm = km ? n - 1 : n;
    const double c_ip = m.cx[n], c_im = m.cx[nim], c_jp = m.cy[n], c_jm = m.cy[njm], c_kp = m.cz[n], c_km = m.cz[nkm];
    const double x0 = xa(n), x_ip = xa(nip), x_im = xa(nim), x_jp = xa(njp), x_jm = xa(njm), x_kp = xa(nkp),
                 x_km = xa(nkm), dg = (double)m.diag[n];
    mem_fence_compiler();
    double v = dg * x0;
    v += ip ? c_ip * x_ip : 0.0;
    v += im ? c_im * x_im : 0.0;
    v += jp ? c_jp * x_jp : 0.0;
    v += jm ? c_jm * x_jm : 0.0;
    v += kp ? c_kp * x_kp : 0.0;
    v += km ? c_km * x_km : 0.0;
    return v;
  } else {
    const long long sx = L.nz, sy = (long long)L.nx * L.nz;
    float c[27];
    double x[27];
    bool ok[27];
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk) {
          const int o = ((dj + 1) * 3 + (di + 1)) * 3 + (dk + 1);
          ok[o] = (unsigned)(j + dj) < (unsigned)L.ny && (unsigned)(i + di) < (unsigned)L.nx &&
                  (unsigned)(k + dk) < (unsigned)L.nz;
          const size_t q = ok[o] ? (size_t)((long long)n + dj * sy + di * sx + dk) : n;
          c[o] = o >= 13 ? L.c14[(size_t)(o - 13) * L.n + n] : L.c14[(size_t)(13 - o) * L.n + q];
          x[o] = xa(q);
        }
    mem_fence_compiler();  // all 54 loads in flight before the first use
    double v = 0;
#pragma unroll
    for (int o = 0; o < 27; ++o) v += ok[o] ? (double)c[o] * x[o] : 0.0;
    return v;
  }
}

// t = x1 + wdinv (b - A x1) with x1 = wdinv b: two Jacobi sweeps from zero
template <bool FINE, class LP, class MA, class LV>
__device__ inline void ph_pre2(const MA& m, const LV& L) {
  auto x1 = [&](size_t q) -> double { return (double)L.wdinv[q] * (double)L.b[q]; };
  for (size_t n = LP::start(); n < L.n; n += LP::stride()) {
    int i, j, k;
    lvl_ijk(L, n, i, j, k);
    L.t[n] = x1(n) + (double)L.wdinv[n] * ((double)L.b[n] - mg_Ax<FINE>(m, L, x1, n, i, j, k));
  }
}

// dst = src + wdinv (b - A src)
template <bool FINE, class LP, class MA, class LV, class R>
__device__ inline void ph_sweep(const MA& m, const LV& L, const R* src, R* dst) {
  auto xs = [&](size_t q) -> double { return src[q]; };
  for (size_t n = LP::start(); n < L.n; n += LP::stride()) {
    int i, j, k;
    lvl_ijk(L, n, i, j, k);
    dst[n] = (double)src[n] + (double)L.wdinv[n] * ((double)L.b[n] - mg_Ax<FINE>(m, L, xs, n, i, j, k));
  }
}

// r = b - A t
template <bool FINE, class LP, class MA, class LV>
__device__ inline void ph_resid(const MA& m, const LV& L) {
  auto xs = [&](size_t q) -> double { return L.t[q]; };
  for (size_t n = LP::start(); n < L.n; n += LP::stride()) {
    int i, j, k;
    lvl_ijk(L, n, i, j, k);
    L.r[n] = (double)L.b[n] - mg_Ax<FINE>(m, L, xs, n, i, j, k);
  }
}

// r = b - A x1 with x1 = wdinv b: one Jacobi sweep from zero and its residual
// (levels >= 1: the restriction stored x1 in x)
template <bool FINE, class LP, class MA, class LV>
__device__ inline void ph_resid1(const MA& m, const LV& L) {
  auto x1 = [&](size_t q) -> double { return L.x[q]; };
  for (size_t n = LP::start(); n < L.n; n += LP::stride()) {
    int i, j, k;
    lvl_ijk(L, n, i, j, k);
    L.r[n] = (double)L.b[n] - mg_Ax<FINE>(m, L, x1, n, i, j, k);
  }
}

__device__ inline double pw1(int f, int C) { return f == 2 * C ? 1.0 : 0.5; }

// b_C = P^T r_F: coarse node I gathers the fine residuals around fine node 2I
// (residuals of fixed fine nodes are zero, so the zeroed rows of P need no test);
// also x_C = wdinv_C b_C, the first sweep from zero on the coarse level.  Fine
// nodes off the mesh are read at a clamped address and add 0.0.
template <class LP, class LV>
__device__ inline void ph_restrict(const LV& F, const LV& C) {
  for (size_t I = LP::start(); I < C.n; I += LP::stride()) {
    int ci, cj, ck;
    lvl_ijk(C, I, ci, cj, ck);
    double r[27];
    bool ok[27];
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk) {
          const int o = ((dj + 1) * 3 + (di + 1)) * 3 + (dk + 1);
          const int fj = 2 * cj + dj, fi = 2 * ci + di, fk = 2 * ck + dk;
          ok[o] = (unsigned)fj < (unsigned)F.ny && (unsigned)fi < (unsigned)F.nx && (unsigned)fk < (unsigned)F.nz;
          r[o] = F.r[ok[o] ? ((size_t)fj * F.nx + fi) * F.nz + fk : 0];
        }
    mem_fence_compiler();
    double v = 0;
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        const double w = (dj == 0 ? 1.0 : 0.5) * (di == 0 ? 1.0 : 0.5);
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk) {
          const int o = ((dj + 1) * 3 + (di + 1)) * 3 + (dk + 1);
          v += ok[o] ? w * (dk == 0 ? 1.0 : 0.5) * r[o] : 0.0;
        }
      }
    C.b[I] = v;
    C.x[I] = (double)C.wdinv[I] * v;
  }
}

// t_F += P e on the unknown nodes (wdinv > 0); with FROM_ZERO the smoothed value
// before the correction is the sweep from zero, t_F = wdinv b + P e (fixed
// nodes 0).  Fine node f takes coarse node f/2 when f is even, (f-1)/2 and
// (f+1)/2 with weight 1/2 each when odd; absent terms add 0.0.
template <class LP, bool FROM_ZERO = false, class LV, class R>
__device__ inline void ph_prolong(const LV& F, const LV& C, const R* e) {
  for (size_t n = LP::start(); n < F.n; n += LP::stride()) {
    const double wd = F.wdinv[n];
    if (wd == 0.0) {
      if (FROM_ZERO) F.t[n] = 0.0;
      continue;
    }
    int i, j, k;
    lvl_ijk(F, n, i, j, k);
    int cj[2] = {j >> 1, (j + 1) >> 1}, ci[2] = {i >> 1, (i + 1) >> 1}, ck[2] = {k >> 1, (k + 1) >> 1};
    bool okj[2] = {cj[0] < C.ny, (j & 1) && cj[1] < C.ny}, oki[2] = {ci[0] < C.nx, (i & 1) && ci[1] < C.nx},
         okk[2] = {ck[0] < C.nz, (k & 1) && ck[1] < C.nz};
    double ev[8];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const bool ok = okj[a] && oki[b] && okk[c];
          ev[(a * 2 + b) * 2 + c] = e[ok ? ((size_t)cj[a] * C.nx + ci[b]) * C.nz + ck[c] : 0];
        }
    mem_fence_compiler();
    double v = 0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const double w = pw1(j, cj[a]) * pw1(i, ci[b]);
#pragma unroll
        for (int c = 0; c < 2; ++c)
          v += (okj[a] && oki[b] && okk[c]) ? w * pw1(k, ck[c]) * ev[(a * 2 + b) * 2 + c] : 0.0;
      }
    if (FROM_ZERO) F.t[n] = wd * (double)F.b[n] + v;
    else F.t[n] = (double)F.t[n] + v;
  }
}

// out = C^-1 in (dense, row major), one wave per row; `wave` and `nwaves` number
// the waves taking part.  A lane sums its columns lane, lane + 64, ... in that
// order; the loads of eight of them are issued before the first multiply, so a
// row costs a few memory round trips instead of one per 64 columns.
template <class MA, class R>
__device__ inline void ph_coarse(const MA& m, const R* in, R* out, size_t wave, size_t nwaves) {
  const int nc = m.nc;
  const int lane = threadIdx.x & 63;
  constexpr int U = 8;
  for (size_t r = wave; r < (size_t)nc; r += nwaves) {
    const R* row = m.cinv + r * nc;
    double v = 0;
    int c = lane;
    for (; c + 64 * (U - 1) < nc; c += 64 * U) {
      double rv[U], xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        rv[u] = row[c + 64 * u];
        xv[u] = in[c + 64 * u];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) v += rv[u] * xv[u];
    }
    for (; c < nc; c += 64) v += (double)row[c] * (double)in[c];
    v = wave_sum(v);
    if (lane == 0) out[r] = v;
  }
}

// Cycle shape: V(2,2) on level 0, V(1,1) on the coarser levels (the same CG
// iteration count on the FinFET meshes, two launches fewer per level).  The
// result of every level >= 1 ends in x (the coarsest is solved into x).
template <class MA>
__device__ __host__ inline const typename MA::real* mg_result(const MA& m, int l) { return m.L[l].x; }

// In a CG iteration the first kernel of the cycle books the residual of the
// iteration that has just updated r (= b of level 0): block 0 writes r.r and the
// iteration count into scal and sets done at the tolerance (the rest of this
// launch still runs; every later kernel returns at once).
template <class MA>
__device__ inline void book_residual(const MA& m) {
  if (!m.part_rr || blockIdx.x != 0) return;
  const double tr = total1(m.part_rr, m.npart);
  if (threadIdx.x == 0) {
    m.scal[0] = tr;
    m.scal[3] += 1.0;
    if (tr <= m.scal[4]) *m.done = 1;
  }
}

template <class AS>
__global__ __launch_bounds__(256) void k_mg_pre2(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  book_residual(m);
  ph_pre2<true, GridLoop>(m, m.L[0]);
}
template <class AS>
__global__ __launch_bounds__(256) void k_mg_resid(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_resid<true, GridLoop>(m, m.L[0]);
}
// levels >= 1: sweep from zero and residual in one pass
template <class AS>
__global__ __launch_bounds__(256) void k_mg_resid1(AS ms_, int l) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_resid1<false, GridLoop>(m, m.L[l]);
}
template <class AS>
__global__ __launch_bounds__(256) void k_mg_restrict(AS ms_, int l) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_restrict<GridLoop>(m.L[l], m.L[l + 1]);
}
template <class AS>
__global__ __launch_bounds__(256) void k_mg_prolong(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_prolong<GridLoop>(m.L[0], m.L[1], mg_result(m, 1));
}
template <class AS>
__global__ __launch_bounds__(256) void k_mg_prolong1(AS ms_, int l) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_prolong<GridLoop, true>(m.L[l], m.L[l + 1], mg_result(m, l + 1));
}
// levels >= 1: one sweep t -> x
template <class AS>
__global__ __launch_bounds__(256) void k_mg_post1(AS ms_, int l) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_sweep<false, GridLoop>(m, m.L[l], m.L[l].t, m.L[l].x);
}
template <class AS>
__global__ __launch_bounds__(256) void k_mg_post0a(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  ph_sweep<true, GridLoop>(m, m.L[0], m.L[0].t, m.L[0].x);
}

// last sweep of level 0 into z, with the partial r.z of the CG (r = b of level 0)
template <class AS>
__global__ __launch_bounds__(256) void k_mg_sweep_last(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  R* z = m.z;
  double* part_rz = m.part_rz;
  if (*m.done) return;
  const auto& L = m.L[0];
  const R* src = L.x;
  auto xs = [&](size_t q) -> double { return src[q]; };
  double rz = 0;
  for (size_t n = GridLoop::start(); n < L.n; n += GridLoop::stride()) {
    int i, j, k;
    lvl_ijk(L, n, i, j, k);
    const double v = (double)src[n] + (double)L.wdinv[n] * ((double)L.b[n] - mg_Ax<true>(m, L, xs, n, i, j, k));
    z[n] = (R)v;
    rz += (double)L.b[n] * v;
  }
  block_reduce1(rz, part_rz);
}

// coarsest level: x = C^-1 b, one wave per row
template <class AS>
__global__ __launch_bounds__(256) void k_mg_coarse(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  if (*m.done) return;
  const auto& L = m.L[m.nl - 1];
  ph_coarse(m, L.b, L.x, ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6, ((size_t)gridDim.x * blockDim.x) >> 6);
}

// single-level hierarchy: z = C^-1 r over the grid, then the partial r.z
template <class AS>
__global__ __launch_bounds__(256) void k_mg_coarse_z(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  R* z = m.z;
  if (*m.done) return;
  book_residual(m);
  ph_coarse(m, m.L[0].b, z, ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6, ((size_t)gridDim.x * blockDim.x) >> 6);
}
template <class AS>
__global__ __launch_bounds__(256) void k_mg_dot_rz(AS ms_) {
  const auto& m = kargs(ms_);
  using R = typename std::decay_t<decltype(m)>::real;
  const R* z = m.z;
  double* part_rz = m.part_rz;
  if (*m.done) return;
  const auto& L = m.L[0];
  double rz = 0;
  for (size_t n = GridLoop::start(); n < L.n; n += GridLoop::stride()) rz += (double)L.b[n] * (double)z[n];
  block_reduce1(rz, part_rz);
}

// ------------------------------------------------------------------ initial population
// init_particle_data (src/particle_init.cpp:112-219) for particle s: position
// uniform in its cell, then select_kstate (Flight::sample_kstate); draws from
// the particle's stream at the reserved step 2^64-1, as the host sampler.
__global__ __launch_bounds__(256) void k_sample_population(SampleArgs a) {
  const TallyDev T{};
  for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < a.n; s += (long long)gridDim.x * blockDim.x) {
    long long lo = 0, hi = a.ngroup - 1;  // last group with start <= s
    while (lo < hi) {
      const long long mid = (lo + hi + 1) >> 1;
      if ((long long)a.start[mid] <= s) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t pk = a.gpk[lo];
    const int i = pk_i(pk), j = pk_j(pk), k = pk_k(pk);
    const MeshDev& M = a.C->M;
    Flight f(a.C->B, M, a.C->F, T);
    const uint64_t key = init_key(a.seed, (uint64_t)s);
    f.rng.init(key, ~0ull);
    double lam = f.Random();
    const double x = lam * M.lx[i] + (1 - lam) * M.lx[i + 1];
    lam = f.Random();
    const double y = lam * M.ly[j] + (1 - lam) * M.ly[j + 1];
    lam = f.Random();
    const double z = lam * M.lz[k] + (1 - lam) * M.lz[k + 1];
    f.par_type = pk_type(pk);
    f.sample_kstate(a.tmp_const);
    a.P.x[s] = x; a.P.y[s] = y; a.P.z[s] = z;
    a.P.kx[s] = f.kx; a.P.ky[s] = f.ky; a.P.kz[s] = f.kz;
    a.P.energy[s] = f.energy; a.P.charge[s] = a.gq[lo];
    a.P.pk[s] = pk; a.P.itet[s] = f.itet; a.P.isym[s] = f.isym; a.P.key[s] = key;
  }
}

// ------------------------------------------------------------------ helpers
__global__ void k_zero_u64(unsigned long long* p, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = 0;
}

// ---------------------------------------------------------------- mixed precision
// r = b - A x in FP64 (0 at the fixed nodes, which k_pcg_fix set to b), its
// FP32 copy, and the partials of r.r and b.b
__global__ __launch_bounds__(256) void k_refine_resid(RefineArgs ra) {
  const PcgArgs& a = ra.a;
  double rr = 0, bb = 0;
  GRID_LOOP(n, a.n) {
    int i, j, k;
    node_ijk(a, n, i, j, k);
    const double r = a.kind[n] != 0 ? 0.0 : a.b[n] - apply_op(a, a.x, n, i, j, k);
    a.r[n] = r;
    ra.r32[n] = (float)r;
    rr += r * r;
    bb += a.b[n] * a.b[n];
  }
  block_reduce3(rr, bb, 0.0, a.part_a);
}
// scal[0] = r.r, scal[2] = b.b (one block)
__global__ __launch_bounds__(256) void k_refine_norms(RefineArgs ra) {
  double t[3];
  total3(ra.a.part_a, ra.a.nblocks, t);
  if (threadIdx.x == 0) {
    ra.a.scal[0] = t[0];
    ra.a.scal[2] = t[1];
  }
}
// x += d (the FP32 correction, in FP64)
__global__ __launch_bounds__(256) void k_refine_add(RefineArgs ra) {
  GRID_LOOP(n, ra.a.n) ra.a.x[n] += (double)ra.d32[n];
}
__global__ void k_zero_f32(float* p, size_t n) {
  GRID_LOOP(q, n) p[q] = 0.0f;
}

// per device of a batch: dst[q] = src[q] (src null: 0) for q < n (readbacks of
// many devices gathered into one block, and the clearing of their counters)
__global__ void k_copy_u64_b(const CopyArgs* t) {
  const CopyArgs& c = t[blockIdx.y];
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < c.n; q += (size_t)gridDim.x * blockDim.x)
    c.dst[q] = !c.src ? 0ull
               : c.src_u32 ? (unsigned long long)((const unsigned*)c.src)[q]
                           : ((const unsigned long long*)c.src)[q];
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

// a.nslow must read zero on entry.  The slow list is short and its length is only
// known on the device: k_fly_slow runs one wave per SIMD (a.wave_slots waves) and
// strides over the list, so no empty workgroups of this large kernel are queued.
hipError_t emc_launch_fly(const FlyArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fly_fast<FlyArgs>, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
  const long long slow_blocks = a.wave_slots > 4 ? a.wave_slots / 4 : 1;
  hipLaunchKernelGGL(k_fly_slow<FlyArgs>, dim3((unsigned)slow_blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_fly_clones(const FlyArgs& a, unsigned long long b, unsigned long long e, hipStream_t s) {
  if (e <= b) return hipSuccess;
  const long long n = (long long)(e - b);
  const long long slots = a.wave_slots > 0 ? a.wave_slots : 1;
  const int lpw = (int)(n <= slots ? 1 : ((n + slots - 1) / slots < 64 ? (n + slots - 1) / slots : 64));
  const long long waves = (n + lpw - 1) / lpw;
  hipLaunchKernelGGL(k_fly_clones, dim3(grid_for(waves * 64, 256)), dim3(256), 0, s, a, b, e, lpw);
  return hipGetLastError();
}
hipError_t emc_launch_merge(const FlyArgs& a, unsigned long long begin, unsigned long long end,
                            const unsigned long long* ndead, unsigned long long* ticket, hipStream_t s) {
  if (end <= begin) return hipSuccess;
  hipLaunchKernelGGL(k_merge, dim3(grid_for((long long)(end - begin), 256)), dim3(256), 0, s, a, begin, end, ndead,
                     ticket);
  return hipGetLastError();
}
// Two clone passes (clones, then the clones they made) and the merge of what
// they flew, with every range read on the device: no host round trip.  ctr[4]
// and ctr[5] must read zero on entry; afterwards ctr[4] .. ctr[5] holds clones
// of a third generation still to fly (almost always empty).
hipError_t emc_launch_clones_dev(const FlyArgs& a, hipStream_t s) {
  const long long blocks = a.wave_slots > 4 ? a.wave_slots / 4 : 1;
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(k_clone_range<FlyArgs>, dim3(1), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_fly_clones_dev<FlyArgs>, dim3((unsigned)blocks), dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(k_clone_range<FlyArgs>, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_merge_dev<FlyArgs>, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}
// sort P into Q by tile; toff[ntile] is the live count afterwards
// One block per tile segment, then blocks striding over the movers outside the
// segments: at most 2048 (the per-particle work is short, and one block per 256
// particles, 12 k blocks at 3 M, costs more in block dispatch).
static unsigned sort_grid(int ntile, long long nappend) {
  return (unsigned)ntile + (unsigned)grid_for(nappend > 0 ? nappend : 1, 256, 2048);
}
hipError_t emc_launch_tile_sort(const SortArgs& a, long long nappend, hipStream_t s) {
  const unsigned g = sort_grid(a.ntile, nappend);
  hipLaunchKernelGGL(k_tile_hist<SortArgs>, dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_tile_scan<SortArgs>, dim3(1), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(k_tile_scatter<SortArgs>, dim3(g), dim3(512), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_deposit(const DepositArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_deposit_tile<DepositArgs>, dim3(a.ntile), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_deposit_gather<DepositArgs>, dim3(grid_for((long long)a.nnode, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_rhs(const RhsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_rhs, dim3(grid_for((long long)a.nnode, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_qc_potential(const QcArgs& a, hipStream_t s) {
  if (a.nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_qc_potential, dim3(grid_for(a.nq, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_field(const FieldArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_field<FieldArgs>, dim3(grid_for((long long)a.ncell, 256, 1024)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_sample(const SampleArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_population, dim3(grid_for(a.n, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_zero(unsigned long long* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for((long long)n, 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}
int emc_pcg_blocks(size_t n) { return grid_for((long long)n, 256, 1024); }

// z = M r: one V(2,2) cycle (m.z receives z, m.part_rz the block partials of
// r.z on the CG grid).  AS is the argument block (one device) or a device table
// of K of them (a batch: gridDim.y = K, the same grids as one device).
template <class SHAPE, class AS>
static void launch_vcycle(const SHAPE& m, AS ms, int nblocks, unsigned K, hipStream_t s) {
  const int nl = m.nl;
  if (nl == 1) {
    hipLaunchKernelGGL(k_mg_coarse_z<AS>, dim3(grid_for((long long)m.nc * 64, 256), K), dim3(256), 0, s, ms);
    hipLaunchKernelGGL(k_mg_dot_rz<AS>, dim3(nblocks, K), dim3(256), 0, s, ms);
    return;
  }
  for (int l = 0; l < nl - 1; ++l) {
    const dim3 g(grid_for((long long)m.L[l].n, 256), K);
    if (l == 0) {
      hipLaunchKernelGGL(k_mg_pre2<AS>, g, dim3(256), 0, s, ms);
      hipLaunchKernelGGL(k_mg_resid<AS>, g, dim3(256), 0, s, ms);
    } else {
      hipLaunchKernelGGL(k_mg_resid1<AS>, g, dim3(256), 0, s, ms, l);
    }
    hipLaunchKernelGGL(k_mg_restrict<AS>, dim3(grid_for((long long)m.L[l + 1].n, 256), K), dim3(256), 0, s, ms, l);
  }
  hipLaunchKernelGGL(k_mg_coarse<AS>, dim3(grid_for((long long)m.nc * 64, 256), K), dim3(256), 0, s, ms);
  for (int l = nl - 2; l >= 0; --l) {
    const dim3 g(grid_for((long long)m.L[l].n, 256), K);
    if (l == 0) {
      hipLaunchKernelGGL(k_mg_prolong<AS>, g, dim3(256), 0, s, ms);
      hipLaunchKernelGGL(k_mg_post0a<AS>, g, dim3(256), 0, s, ms);
      hipLaunchKernelGGL(k_mg_sweep_last<AS>, dim3(nblocks, K), dim3(256), 0, s, ms);
    } else {
      hipLaunchKernelGGL(k_mg_prolong1<AS>, g, dim3(256), 0, s, ms, l);
      hipLaunchKernelGGL(k_mg_post1<AS>, g, dim3(256), 0, s, ms, l);
    }
  }
}

// the cycle's argument block of the solve's start (r.z partials into part_rz0,
// no residual booking) and of CG iteration `it` (r.z into the partial array the
// next iteration reads as rz_new; the first kernel books the residual)
// the cycle's output: the CG's z, or z32 for the FP32 cycle of the FP64 CG
template <class RM, class RA>
static RM* cycle_out(const PcgArgsT<RA>& a) {
  if constexpr (std::is_same<RA, RM>::value) return a.z;
  else return const_cast<float*>(a.z32);
}
template <class RA, class RM>
static MgArgsT<RM> mg_start_args(const PcgArgsT<RA>& a, const MgArgsT<RM>& m) {
  MgArgsT<RM> m0 = m;
  m0.part_rr = nullptr;
  m0.z = cycle_out<RM>(a);
  m0.part_rz = a.part_rz0;
  return m0;
}
template <class RA, class RM>
static MgArgsT<RM> mg_iter_args(const PcgArgsT<RA>& a, const MgArgsT<RM>& m, int it) {
  MgArgsT<RM> mi = m;
  mi.part_rr = a.part_rr;
  mi.scal = a.scal;
  mi.npart = a.nblocks;
  mi.z = cycle_out<RM>(a);
  mi.part_rz = (it & 1) ? a.part_rz0 : a.part_rz1;
  return mi;
}

MgArgs emc_mg_start_args(const PcgArgs& a, const MgArgs& m) { return mg_start_args(a, m); }
MgArgs emc_mg_iter_args(const PcgArgs& a, const MgArgs& m, int it) { return mg_iter_args(a, m, it); }
MgArgsT<float> emc_mg_start_args(const PcgArgsT<float>& a, const MgArgsT<float>& m) { return mg_start_args(a, m); }
MgArgsT<float> emc_mg_iter_args(const PcgArgsT<float>& a, const MgArgsT<float>& m, int it) {
  return mg_iter_args(a, m, it);
}
MgArgsT<float> emc_mg_start_args(const PcgArgs& a, const MgArgsT<float>& m) { return mg_start_args(a, m); }
MgArgsT<float> emc_mg_iter_args(const PcgArgs& a, const MgArgsT<float>& m, int it) { return mg_iter_args(a, m, it); }

template <class RA, class RM>
static hipError_t launch_pcg_start(const PcgArgsT<RA>& a, const MgArgsT<RM>& m, hipStream_t s) {
  using PA = PcgArgsT<RA>;
  using MA = MgArgsT<RM>;
  const int g = a.nblocks;
  hipLaunchKernelGGL(k_pcg_fix<PA>, dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_pcg_resid<PA>, dim3(g), dim3(256), 0, s, a);
  const MA m0 = mg_start_args(a, m);
  launch_vcycle<MA, MA>(m0, m0, g, 1, s);
  hipLaunchKernelGGL(k_pcg_begin<PA>, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t emc_launch_pcg_start(const PcgArgs& a, const MgArgs& m, hipStream_t s) { return launch_pcg_start(a, m, s); }
hipError_t emc_launch_pcg_start(const PcgArgsT<float>& a, const MgArgsT<float>& m, hipStream_t s) {
  return launch_pcg_start(a, m, s);
}
hipError_t emc_launch_pcg_start(const PcgArgs& a, const MgArgsT<float>& m, hipStream_t s) { return launch_pcg_start(a, m, s); }

// iteration `it`: r.z of the previous iteration is in part_rz0 (even it) or
// part_rz1; the search direction alternates between p (even it) and p2
template <class RA, class RM>
static hipError_t launch_pcg_iter(const PcgArgsT<RA>& a, const MgArgsT<RM>& m, int it, hipStream_t s) {
  using PA = PcgArgsT<RA>;
  using MA = MgArgsT<RM>;
  const int g = a.nblocks;
  hipLaunchKernelGGL(k_pcg_spmvp<PA>, dim3(g), dim3(256), 0, s, a, it);
  hipLaunchKernelGGL(k_pcg_update<PA>, dim3(g), dim3(256), 0, s, a, it);
  const MA mi = mg_iter_args(a, m, it);
  launch_vcycle<MA, MA>(mi, mi, g, 1, s);
  return hipGetLastError();
}
hipError_t emc_launch_pcg_iter(const PcgArgs& a, const MgArgs& m, int it, hipStream_t s) {
  return launch_pcg_iter(a, m, it, s);
}
hipError_t emc_launch_pcg_iter(const PcgArgsT<float>& a, const MgArgsT<float>& m, int it, hipStream_t s) {
  return launch_pcg_iter(a, m, it, s);
}
hipError_t emc_launch_pcg_iter(const PcgArgs& a, const MgArgsT<float>& m, int it, hipStream_t s) {
  return launch_pcg_iter(a, m, it, s);
}

// ---- iterative refinement of the mixed-precision solve
hipError_t emc_launch_refine_resid(const RefineArgs& r, hipStream_t s) {
  const int g = r.a.nblocks;
  hipLaunchKernelGGL(k_pcg_fix<PcgArgs>, dim3(g), dim3(256), 0, s, r.a);
  hipLaunchKernelGGL(k_refine_resid, dim3(g), dim3(256), 0, s, r);
  hipLaunchKernelGGL(k_refine_norms, dim3(1), dim3(256), 0, s, r);
  return hipGetLastError();
}
hipError_t emc_launch_refine_add(const RefineArgs& r, hipStream_t s) {
  hipLaunchKernelGGL(k_refine_add, dim3(grid_for((long long)r.a.n, 256, 2048)), dim3(256), 0, s, r);
  return hipGetLastError();
}
hipError_t emc_launch_zero_f32(float* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_f32, dim3(grid_for((long long)n, 256, 2048)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// ---- a batch of K devices sharing the mesh: one launch per kernel for all of
// them (gridDim.y = K), each device's blocks identical to its own launch.
hipError_t emc_launch_pcg_start_b(const BatchTables& t, const MgArgs& shape, int nblocks, hipStream_t s) {
  const unsigned K = (unsigned)t.K;
  hipLaunchKernelGGL(k_pcg_fix<const PcgArgs*>, dim3(nblocks, K), dim3(256), 0, s, t.pcg);
  hipLaunchKernelGGL(k_pcg_resid<const PcgArgs*>, dim3(nblocks, K), dim3(256), 0, s, t.pcg);
  if (t.fmg_start) launch_vcycle<MgArgs, const MgArgsT<float>*>(shape, t.fmg_start, nblocks, K, s);
  else launch_vcycle<MgArgs, const MgArgs*>(shape, t.mg_start, nblocks, K, s);
  hipLaunchKernelGGL(k_pcg_begin<const PcgArgs*>, dim3(1, K), dim3(256), 0, s, t.pcg);
  return hipGetLastError();
}
hipError_t emc_launch_pcg_iter_b(const BatchTables& t, const MgArgs& shape, int nblocks, int it, hipStream_t s) {
  const unsigned K = (unsigned)t.K;
  hipLaunchKernelGGL(k_pcg_spmvp<const PcgArgs*>, dim3(nblocks, K), dim3(256), 0, s, t.pcg, it);
  hipLaunchKernelGGL(k_pcg_update<const PcgArgs*>, dim3(nblocks, K), dim3(256), 0, s, t.pcg, it);
  if (t.fmg_start) launch_vcycle<MgArgs, const MgArgsT<float>*>(shape, (it & 1) ? t.fmg_odd : t.fmg_even, nblocks, K, s);
  else launch_vcycle<MgArgs, const MgArgs*>(shape, (it & 1) ? t.mg_odd : t.mg_even, nblocks, K, s);
  return hipGetLastError();
}
hipError_t emc_launch_tile_sort_b(const BatchTables& t, int ntile, long long nappend_max, hipStream_t s) {
  const unsigned K = (unsigned)t.K;
  const unsigned g = sort_grid(ntile, nappend_max);
  hipLaunchKernelGGL(k_tile_hist<const SortArgs*>, dim3(g, K), dim3(256), 0, s, t.sort);
  hipLaunchKernelGGL(k_tile_scan<const SortArgs*>, dim3(1, K), dim3(1024), 0, s, t.sort);
  hipLaunchKernelGGL(k_tile_scatter<const SortArgs*>, dim3(g, K), dim3(512), 0, s, t.sort);
  return hipGetLastError();
}
hipError_t emc_launch_deposit_b(const BatchTables& t, int ntile, size_t nnode, hipStream_t s) {
  const unsigned K = (unsigned)t.K;
  hipLaunchKernelGGL(k_deposit_tile<const DepositArgs*>, dim3(ntile, K), dim3(256), 0, s, t.dep);
  hipLaunchKernelGGL(k_deposit_gather<const DepositArgs*>, dim3(grid_for((long long)nnode, 256), K), dim3(256), 0, s,
                     t.dep);
  return hipGetLastError();
}
hipError_t emc_launch_field_b(const BatchTables& t, size_t ncell, hipStream_t s) {
  hipLaunchKernelGGL(k_field<const FieldArgs*>, dim3(grid_for((long long)ncell, 256, 1024), (unsigned)t.K), dim3(256), 0,
                     s, t.field);
  return hipGetLastError();
}
hipError_t emc_launch_fly_b(const BatchTables& t, long long nmax, long long wave_slots, hipStream_t s) {
  const unsigned K = (unsigned)t.K;
  if (nmax > 0) hipLaunchKernelGGL(k_fly_fast<const FlyArgs*>, dim3(grid_for(nmax, 256), K), dim3(256), 0, s, t.fly);
  const long long blocks = wave_slots > 4 ? wave_slots / 4 : 1;
  hipLaunchKernelGGL(k_fly_slow<const FlyArgs*>, dim3((unsigned)blocks, K), dim3(256), 0, s, t.fly);
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(k_clone_range<const FlyArgs*>, dim3(1, K), dim3(64), 0, s, t.fly);
    hipLaunchKernelGGL(k_fly_clones_dev<const FlyArgs*>, dim3((unsigned)blocks, K), dim3(256), 0, s, t.fly);
  }
  hipLaunchKernelGGL(k_clone_range<const FlyArgs*>, dim3(1, K), dim3(64), 0, s, t.fly);
  hipLaunchKernelGGL(k_merge_dev<const FlyArgs*>, dim3((unsigned)blocks, K), dim3(256), 0, s, t.fly);
  return hipGetLastError();
}
hipError_t emc_launch_copy_b(const CopyArgs* t, int K, size_t nwords, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_u64_b, dim3(grid_for((long long)nwords, 256, 64), (unsigned)K), dim3(256), 0, s, t);
  return hipGetLastError();
}
