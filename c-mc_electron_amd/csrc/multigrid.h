// Geometric multigrid hierarchy of the eliminated Poisson operator (host side).
//
// The reference factorises the Poisson matrix once with KLU (src/PoissonSolver.cpp:
// 48-78) and back-solves every step.  This build solves the same SPD system (contact
// rows eliminated) with CG preconditioned by one multigrid cycle, V(2,2) on
// level 0 and V(1,1) on the coarser levels:
//   - level 0 is the 7-point operator (diag, cx, cy, cz) of the node grid;
//   - level l+1 has floor(n_l/2)+1 nodes per axis, coarse node I sitting on fine
//     node 2I; P is trilinear interpolation with the rows of fixed (Dirichlet or
//     identity) fine nodes zeroed, and A_{l+1} = P^T A_l P (Galerkin, 27 points);
//     coarse rows that come out empty are fixed nodes (identity);
//   - smoothing is damped Jacobi, omega_l = 4 / (3 lambda_max(D^-1 A_l)), with
//     lambda_max from a power iteration;
//   - the coarsest level (at most `min_nodes` nodes) is solved with its dense inverse.
// Everything here is fixed for a run: it depends on the mesh and permittivities only.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace emc {

struct Device;

// offset index of a 27-point stencil entry: o = ((dj+1)*3 + (di+1))*3 + (dk+1)
inline int st27(int di, int dj, int dk) { return ((dj + 1) * 3 + (di + 1)) * 3 + (dk + 1); }

struct MgLevel {
  int nx = 0, ny = 0, nz = 0;  // nodes per axis (i, j, k); node n = (j*nx + i)*nz + k
  size_t n = 0;
  std::vector<double> a27;     // [27][n] coefficients, level >= 1 only
  std::vector<double> wdinv;   // omega / diag, 0 at fixed nodes
  std::vector<uint8_t> fixed;  // 1: Dirichlet or identity node
  double omega = 0, lambda_max = 0;
};

struct MgHierarchy {
  std::vector<MgLevel> lv;
  std::vector<double> cinv;  // dense inverse of the coarsest operator, [n*n] row major
  double coarse_shift = 0.0; // diagonal shift that made it factorise (0: none needed)
};

// Builds the hierarchy from the device's eliminated 7-point operator; `threads`
// host threads compute the Galerkin products.  Returns "" or an error message.
std::string mg_build(const Device& d, MgHierarchy& h, int min_nodes, int threads);

// y = A_l x on the host (used by the builder and by tests)
void mg_apply(const Device& d, const MgHierarchy& h, int l, const double* x, double* y);

}  // namespace emc
