// Quantum-correction effective potential (src/mcmodel.cpp:2337-2405, 2406-2662):
// the electrons' potential at a node of the quantum region is the Poisson
// potential (plus the barrier Eb in oxide) smoothed by a Gaussian of widths
// qc_*theta, each source node weighted by the Gaussian's integral over its
// control box:
//
//   V_eff(p) = 1/8 sum_s (V(s) + Eb(s)) (erf(z2)-erf(z1)) (erf(y2)-erf(y1)) (erf(x2)-erf(x1))
//
// over the silicon and oxide nodes s of a window of +-qc_*ratio*qc_*theta around
// p.  The erf differences depend on the geometry only: they are tabulated once
// per axis, for each (target index, source index) and for the four box shapes a
// source can have on that axis (its low and high side cut by the region edge or
// not), so the per-step work is the weighted sum alone.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "device.h"

namespace emc {

// init_p_mat node flags (include/defs.h:129-137)
constexpr uint16_t QN_SILICON = 1, QN_OXIDE = 2, QN_UP = 4, QN_DOWN = 8, QN_FRONT = 16, QN_BACK = 32,
                   QN_LEFT = 64, QN_RIGHT = 128, QN_QUANTUM = 256;

struct QcTables {
  std::vector<uint16_t> flag;     // [nnode]
  std::vector<int> lo[3], hi[3];  // per axis (x=i, y=j, z=k): source window [lo, hi] of each target index
  int wlen[3] = {0, 0, 0};        // longest window per axis
  // per axis: w[a][(t * wlen + (s - lo[t])) * 4 + v] = erf(hi edge) - erf(lo edge) of
  // source s seen from target t, box shape v (bit 0: low side at the region edge,
  // bit 1: high side at the region edge)
  std::vector<double> w[3];
  std::vector<int> qnodes;        // nodes of the quantum region, ascending
  double eb = 0;                  // normalised barrier, added at oxide-only nodes as -Eb
};

// init_p_mat, init_effective_potential and init_ep_range for the device.
std::string build_qc_tables(const Device& d, QcTables& t);

// Box shape of node flags f on axis a.
inline int qc_shape(uint16_t f, int a) {
  const uint16_t lo = a == 0 ? QN_UP : a == 1 ? QN_LEFT : QN_FRONT;
  const uint16_t hi = a == 0 ? QN_DOWN : a == 1 ? QN_RIGHT : QN_BACK;
  return ((f & lo) ? 1 : 0) | ((f & hi) ? 2 : 0);
}

}  // namespace emc
