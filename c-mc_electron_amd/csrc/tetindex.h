// Point-location index over the tetrahedra of each band (host-side builder).
//
// Getisym (src/ParticleUtils.cpp:395-625) finds the tetrahedron holding a k-point by
// scanning energy buckets of the tetrahedron lists in list order (tlists bin, then
// tlist bins ie, ie+1, ie-1), and as a last resort every tetrahedron in index order.
// The buckets hold up to a few thousand tetrahedra, so a scan costs thousands of
// dependent loads.  The index gives the same answer from a few dozen candidates:
//   - one octree per band over the boxes of the grown tetrahedra (every point
//     Flight::contains accepts lies in its tetrahedron's box); a leaf lists every
//     tetrahedron of the band whose box meets the leaf's closed box, so the leaf
//     holding a point lists every tetrahedron that can hold it;
//   - leaf entries run in descending tetrahedron index, which is the order of every
//     bucket (checked when the index is built), and carry the bucket range of their
//     tetrahedron: a tetrahedron is in bucket (bin, band) of tlist exactly when
//     lo <= bin <= hi (contiguity checked when the index is built), likewise for tlists;
//   - the first holding entry of a bucket is the bucket scan's answer, the last
//     holding entry the full scan's.
// When a check fails the index is not built and Getisym scans the lists.
#pragma once
#include <string>
#include <vector>

#include "tables.h"

namespace emc {

struct Band;

// Axis-aligned box of a grown tetrahedron (host side; infinite when degenerate).
struct TetBox {
  double lo[3], hi[3];
};

struct TetIndex {
  std::vector<TetIdxRoot> roots;  // [MNB]
  std::vector<TetIdxNode> nodes;
  std::vector<TetIdxEntry> ent;
  int max_depth = 0, max_leaf = 0;
};

// Builds the index from the band tables and the grown-tetrahedron boxes.
// Returns "" or the reason the index cannot reproduce the list scans.
std::string build_tet_index(const Band& b, const std::vector<TetBox>& box, TetIndex& ix, int leaf_max = 16,
                            int depth_max = 14);

}  // namespace emc
