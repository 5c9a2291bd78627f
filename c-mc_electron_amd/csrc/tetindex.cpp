// Point-location index over the tetrahedra of each band: builder and checks.
#include "tetindex.h"

#include <algorithm>
#include <climits>
#include <cmath>

#include "band.h"
#include "emc_const.h"

namespace emc {

namespace {

// Bucket range of every tetrahedron in one list family; checks that each bucket
// runs in descending tetrahedron index and that each tetrahedron's buckets of its
// own band are one contiguous run of bins.
std::string bucket_ranges(const Band& b, const std::vector<int>& cnt, const std::vector<int>& start,
                          const std::vector<int>& list, int nbin, const char* name, std::vector<int16_t>& lo,
                          std::vector<int16_t>& hi) {
  const int nt = b.nt;
  std::vector<int> l(nt, INT_MAX), h(nt, -1), n(nt, 0);
  if ((int)cnt.size() < nbin * MNB || (int)start.size() < nbin * MNB) return std::string(name) + ": bucket table too short";
  for (int ie = 0; ie < nbin; ++ie)
    for (int band = 0; band < MNB; ++band) {
      const int p0 = start[ie * MNB + band], nn = cnt[ie * MNB + band];
      if (nn == 0) continue;
      if (p0 < 0 || nn < 0 || (size_t)p0 + (size_t)nn > list.size()) return std::string(name) + ": bucket out of range";
      int prev = INT_MAX;
      for (int q = p0; q < p0 + nn; ++q) {
        const int t = list[q];
        if (t < 0 || t >= nt) return std::string(name) + ": entry out of range";
        if (t >= prev) return std::string(name) + ": bucket not in descending tetrahedron order";
        prev = t;
        if (b.ibt[t] != band) continue;  // never accepted by the scan
        l[t] = std::min(l[t], ie);
        h[t] = std::max(h[t], ie);
        ++n[t];
      }
    }
  lo.assign(nt, 1);
  hi.assign(nt, 0);
  for (int t = 0; t < nt; ++t) {
    if (n[t] == 0) continue;
    if (n[t] != h[t] - l[t] + 1) return std::string(name) + ": buckets of a tetrahedron not contiguous";
    lo[t] = (int16_t)l[t];
    hi[t] = (int16_t)h[t];
  }
  return "";
}

struct Builder {
  const Band& b;
  const std::vector<TetBox>& box;
  const std::vector<int16_t>&lo, &hi, &slo, &shi;
  TetIndex& ix;
  int leaf_max, depth_max;

  // false only when the node box provably misses the grown tetrahedron: disjoint
  // from its box, or wholly outside one of its faces (n.x >= d - MY_ZERO fails at
  // every corner of the node box, by more than a rounding margin)
  bool meets(int it, const double* blo, const double* bhi) const {
    const TetBox& t = box[it];
    for (int a = 0; a < 3; ++a)
      if (t.hi[a] < blo[a] || t.lo[a] > bhi[a]) return false;
    for (int f = 0; f < 4; ++f) {
      double top = 0, mag = std::fabs(b.dn(3, f, it));
      for (int a = 0; a < 3; ++a) {
        const double n = b.dn(a, f, it);
        top += n >= 0 ? n * bhi[a] : n * blo[a];
        mag += std::fabs(n) * std::max(std::fabs(blo[a]), std::fabs(bhi[a]));
      }
      if (top - b.dn(3, f, it) + MY_ZERO < -1e-9 * mag - 1e-12) return false;
    }
    return true;
  }

  TetIdxEntry entry(int t) const {
    const TetBox& x = box[t];
    TetIdxEntry e{};
    for (int a = 0; a < 3; ++a) {
      float l = (float)x.lo[a], h = (float)x.hi[a];
      if ((double)l > x.lo[a]) l = std::nextafter(l, -INFINITY);
      if ((double)h < x.hi[a]) h = std::nextafter(h, INFINITY);
      e.lo[a] = l;
      e.hi[a] = h;
    }
    e.t = t;
    e.bins = (uint32_t)(lo[t] & 511) | (uint32_t)(hi[t] & 511) << 9 | (uint32_t)(slo[t] & 63) << 18 |
             (uint32_t)(shi[t] & 63) << 24;
    return e;
  }

  // fills node `k` for the box blo..bhi and the candidates `c` (descending index)
  void build(int k, const double* blo, const double* bhi, const std::vector<int>& c, int depth) {
    ix.max_depth = std::max(ix.max_depth, depth);
    if ((int)c.size() > leaf_max && depth < depth_max) {
      double mid[3];
      for (int a = 0; a < 3; ++a) mid[a] = 0.5 * (blo[a] + bhi[a]);
      std::vector<int> sub[8];
      double clo[8][3], chi[8][3];
      size_t total = 0;
      for (int ch = 0; ch < 8; ++ch) {
        for (int a = 0; a < 3; ++a) {
          const bool up = (ch >> a) & 1;
          clo[ch][a] = up ? mid[a] : blo[a];
          chi[ch][a] = up ? bhi[a] : mid[a];
        }
        for (int t : c)
          if (meets(t, clo[ch], chi[ch])) sub[ch].push_back(t);
        total += sub[ch].size();
      }
      // split unless most boxes would land in several children (boxes larger than
      // the node stay in every child, and splitting again only copies them)
      if (total <= 3 * c.size()) {
        const int first = (int)ix.nodes.size();
        ix.nodes.resize(ix.nodes.size() + 8);
        ix.nodes[k] = TetIdxNode{first, -1};
        for (int ch = 0; ch < 8; ++ch) build(first + ch, clo[ch], chi[ch], sub[ch], depth + 1);
        return;
      }
    }
    ix.nodes[k] = TetIdxNode{(int)ix.ent.size(), (int)c.size()};
    ix.max_leaf = std::max(ix.max_leaf, (int)c.size());
    for (int t : c) ix.ent.push_back(entry(t));
  }
};

}  // namespace

std::string build_tet_index(const Band& b, const std::vector<TetBox>& box, TetIndex& ix, int leaf_max,
                            int depth_max) {
  static_assert(MWLE <= 512 && MWLES <= 64, "bucket bins must fit the packed ranges");
  ix = TetIndex{};
  if ((int)box.size() != b.nt || (int)b.ibt.size() != b.nt) return "box or band table size";
  std::vector<int16_t> lo, hi, slo, shi;
  std::string e = bucket_ranges(b, b.ntlist, b.ptlist, b.tlist, MWLE, "tlist", lo, hi);
  if (e.empty()) e = bucket_ranges(b, b.ntlists, b.ptlists, b.tlists, MWLES, "tlists", slo, shi);
  if (!e.empty()) return e;
  for (int t = 0; t < b.nt; ++t) {
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(box[t].lo[a]) || !std::isfinite(box[t].hi[a]) || std::fabs(box[t].lo[a]) > 1e30 ||
          std::fabs(box[t].hi[a]) > 1e30)
        return "degenerate tetrahedron";
    if (b.ibt[t] < 0 || b.ibt[t] >= MNB) return "band index out of range";
  }
  Builder bl{b, box, lo, hi, slo, shi, ix, leaf_max, depth_max};
  ix.roots.assign(MNB, TetIdxRoot{});
  for (int band = 0; band < MNB; ++band) {
    std::vector<int> c;
    for (int t = b.nt - 1; t >= 0; --t)
      if (b.ibt[t] == band) c.push_back(t);
    TetIdxRoot& R = ix.roots[band];
    for (int a = 0; a < 3; ++a) { R.lo[a] = 0.0; R.hi[a] = 0.0; }
    if (!c.empty()) {
      for (int a = 0; a < 3; ++a) { R.lo[a] = HUGE_VAL; R.hi[a] = -HUGE_VAL; }
      for (int t : c)
        for (int a = 0; a < 3; ++a) {
          R.lo[a] = std::min(R.lo[a], box[t].lo[a]);
          R.hi[a] = std::max(R.hi[a], box[t].hi[a]);
        }
      for (int a = 0; a < 3; ++a) {
        const double w = 1e-6 * (R.hi[a] - R.lo[a]) + 1e-9 * (std::fabs(R.lo[a]) + std::fabs(R.hi[a])) + 1e-300;
        R.lo[a] -= w;
        R.hi[a] += w;
      }
    }
    R.node = (int)ix.nodes.size();
    ix.nodes.emplace_back();
    bl.build(R.node, R.lo, R.hi, c, 0);
  }
  return "";
}

}  // namespace emc
