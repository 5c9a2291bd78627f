// Tables of the quantum-correction effective potential (src/mcmodel.cpp:2406-2662).
#include "qc.h"

#include <algorithm>
#include <cmath>

namespace emc {

namespace {

constexpr double SQRT2 = 1.4142135623730951;  // sqrt(2), correctly rounded

// init_effective_potential's window of one axis (src/mcmodel.cpp:2584-2625): for
// node p, the first node at or after l[p] - r (one before it when there is one)
// through the first node after l[p] + r (the last node when there is none).  The
// searches resume where the previous node's ended, the lower one after its
// step back, as the reference's iterators do.
void axis_window(const std::vector<double>& l, double r, std::vector<int>& lo, std::vector<int>& hi) {
  const int n = (int)l.size();
  lo.resize(n);
  hi.resize(n);
  int a = 0, b = 0;
  for (int p = 0; p < n; ++p) {
    a = (int)(std::lower_bound(l.begin() + a, l.begin() + std::max(a, p), l[p] - r) - l.begin());
    b = (int)(std::upper_bound(l.begin() + b, l.end(), l[p] + r) - l.begin());
    if (a != 0) --a;
    if (b == n) --b;
    lo[p] = a;
    hi[p] = b;
  }
}

}  // namespace

std::string build_qc_tables(const Device& d, QcTables& t) {
  const int nx = d.pnx, ny = d.pny, nz = d.pnz;
  const size_t NN = d.nnode;
  t.flag.assign(NN, 0);
  // init_p_mat (src/mcmodel.cpp:2495-2582): the 8 corners of each cell
  for (int i = 0; i < d.cnx; ++i)
    for (int k = 0; k < d.cnz; ++k)
      for (int j = 0; j < d.cny; ++j) {
        const size_t c = d.cell(i, j, k);
        uint16_t f = 0;
        if (d.qcregion[c]) f |= QN_QUANTUM;
        if (d.material[c] == SILICON) f |= QN_SILICON;
        if (d.material[c] == OXIDE) f |= QN_OXIDE;
        if (!f) continue;
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) t.flag[d.node(i + di, j + dj, k + dk)] |= f;
      }
  std::vector<uint16_t> mat(t.flag);
  auto solid = [&](int i, int j, int k) { return (mat[d.node(i, j, k)] & (QN_OXIDE | QN_SILICON)) != 0; };
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        uint16_t& f = t.flag[d.node(i, j, k)];
        if (i == 0 || !solid(i - 1, j, k)) f |= QN_UP;
        if (i == nx - 1 || !solid(i + 1, j, k)) f |= QN_DOWN;
        if (k == 0 || !solid(i, j, k - 1)) f |= QN_FRONT;
        if (k == nz - 1 || !solid(i, j, k + 1)) f |= QN_BACK;
        if (j == 0 || !solid(i, j - 1, k)) f |= QN_LEFT;
        if (j == ny - 1 || !solid(i, j + 1, k)) f |= QN_RIGHT;
      }
  // init_effective_potential and init_ep_range (src/mcmodel.cpp:2406-2493, 2584-2625)
  const std::vector<double>* L[3] = {&d.lx, &d.ly, &d.lz};
  const std::vector<double>* D[3] = {&d.dx, &d.dy, &d.dz};
  for (int a = 0; a < 3; ++a) {
    const std::vector<double>& l = *L[a];
    const std::vector<double>& dl = *D[a];
    const double th = d.qc_theta[a], r = d.qc_ratio[a] * th, s2t = SQRT2 * th;
    axis_window(l, r, t.lo[a], t.hi[a]);
    const int n = (int)l.size();
    int wl = 0;
    for (int p = 0; p < n; ++p) wl = std::max(wl, t.hi[a][p] - t.lo[a][p] + 1);
    t.wlen[a] = wl;
    t.w[a].assign((size_t)n * wl * 4, 0.0);
    for (int p = 0; p < n; ++p) {
      const double q = l[p] / s2t;
      for (int s = t.lo[a][p]; s <= t.hi[a][p]; ++s)
        for (int v = 0; v < 4; ++v) {
          // the low side of the box is cut by the region edge at every first
          // node and the high side at every last node: the other shapes do not
          // occur there
          if ((s == 0 && !(v & 1)) || (s == n - 1 && !(v & 2))) continue;
          const double e1 = (v & 1) ? l[s] - r : l[s] - 0.5 * dl[s - 1];
          const double e2 = (v & 2) ? l[s] + r : l[s] + 0.5 * dl[s];
          const double x1 = e1 / s2t, x2 = e2 / s2t;
          t.w[a][((size_t)p * wl + (s - t.lo[a][p])) * 4 + v] = std::erf(x2 - q) - std::erf(x1 - q);
        }
    }
  }
  t.qnodes.clear();
  for (size_t n = 0; n < NN; ++n)
    if (t.flag[n] & QN_QUANTUM) t.qnodes.push_back((int)n);
  t.eb = d.qc_eb;
  (void)nz;
  return "";
}

}  // namespace emc
