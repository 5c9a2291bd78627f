// Geometric multigrid hierarchy of the eliminated Poisson operator (host side).
#include "multigrid.h"

#include <algorithm>
#include <cmath>
#include <thread>

#include "device.h"

namespace emc {

namespace {

// interpolation weight of coarse index C at fine index f along one axis
inline double pw(int f, int C) {
  const int d = f - 2 * C;
  return d == 0 ? 1.0 : (d == 1 || d == -1) ? 0.5 : 0.0;
}

// Entry A_l(n, n + offset) where the offset is (di, dj, dk) in {-1,0,1}^3.
struct Op {
  const Device& d;
  const MgLevel& L;
  bool fine;  // level 0: 7-point arrays of the device
  double at(size_t n, int i, int j, int k, int di, int dj, int dk) const {
    if (!fine) return L.a27[(size_t)st27(di, dj, dk) * L.n + n];
    const int nz = L.nz, sx = L.nz, sy = L.nx * L.nz;
    const int s = std::abs(di) + std::abs(dj) + std::abs(dk);
    if (s == 0) return d.diag[n];
    if (s != 1) return 0.0;
    if (di == 1) return i + 1 < L.nx ? d.cx[n] : 0.0;
    if (di == -1) return i > 0 ? d.cx[n - sx] : 0.0;
    if (dj == 1) return j + 1 < L.ny ? d.cy[n] : 0.0;
    if (dj == -1) return j > 0 ? d.cy[n - sy] : 0.0;
    if (dk == 1) return k + 1 < nz ? d.cz[n] : 0.0;
    return k > 0 ? d.cz[n - 1] : 0.0;
  }
};

void galerkin(const Device& d, const MgLevel& F, bool fine, MgLevel& Cl, int threads) {
  const Op op{d, F, fine};
  const size_t NC = Cl.n;
  Cl.a27.assign(27 * NC, 0.0);
  auto work = [&](int j0, int j1) {
    for (int cj = j0; cj < j1; ++cj)
      for (int ci = 0; ci < Cl.nx; ++ci)
        for (int ck = 0; ck < Cl.nz; ++ck) {
          const size_t I = ((size_t)cj * Cl.nx + ci) * Cl.nz + ck;
          double acc[27] = {0};
          // fine nodes i with P(i, I) != 0
          for (int fj = 2 * cj - 1; fj <= 2 * cj + 1; ++fj) {
            if (fj < 0 || fj >= F.ny) continue;
            for (int fi = 2 * ci - 1; fi <= 2 * ci + 1; ++fi) {
              if (fi < 0 || fi >= F.nx) continue;
              for (int fk = 2 * ck - 1; fk <= 2 * ck + 1; ++fk) {
                if (fk < 0 || fk >= F.nz) continue;
                const size_t fn = ((size_t)fj * F.nx + fi) * F.nz + fk;
                if (F.fixed[fn]) continue;
                const double wi = pw(fi, ci) * pw(fj, cj) * pw(fk, ck);
                // neighbours j of fine node i
                for (int dj = -1; dj <= 1; ++dj)
                  for (int di = -1; di <= 1; ++di)
                    for (int dk = -1; dk <= 1; ++dk) {
                      const int gi = fi + di, gj = fj + dj, gk = fk + dk;
                      if (gi < 0 || gi >= F.nx || gj < 0 || gj >= F.ny || gk < 0 || gk >= F.nz) continue;
                      const size_t gn = ((size_t)gj * F.nx + gi) * F.nz + gk;
                      if (F.fixed[gn]) continue;
                      const double a = op.at(fn, fi, fj, fk, di, dj, dk);
                      if (a == 0.0) continue;
                      // coarse J with P(g, J) != 0; |J - I| <= 1 on every axis
                      for (int Jj = gj / 2; Jj <= (gj + 1) / 2; ++Jj) {
                        if (Jj >= Cl.ny) continue;
                        const double wj = pw(gj, Jj);
                        for (int Ji = gi / 2; Ji <= (gi + 1) / 2; ++Ji) {
                          if (Ji >= Cl.nx) continue;
                          const double wij = wj * pw(gi, Ji);
                          for (int Jk = gk / 2; Jk <= (gk + 1) / 2; ++Jk) {
                            if (Jk >= Cl.nz) continue;
                            const double w = wij * pw(gk, Jk);
                            acc[st27(Ji - ci, Jj - cj, Jk - ck)] += wi * a * w;
                          }
                        }
                      }
                    }
              }
            }
          }
          for (int o = 0; o < 27; ++o) Cl.a27[(size_t)o * NC + I] = acc[o];
        }
  };
  const int nt = std::max(1, std::min(threads, Cl.ny));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t) pool.emplace_back(work, Cl.ny * t / nt, Cl.ny * (t + 1) / nt);
  for (auto& th : pool) th.join();
  // rows with no unknown fine node in their support become identity rows
  Cl.fixed.assign(NC, 0);
  for (size_t I = 0; I < NC; ++I) {
    bool any = false;
    for (int o = 0; o < 27 && !any; ++o) any = Cl.a27[(size_t)o * NC + I] != 0.0;
    if (!any) {
      Cl.fixed[I] = 1;
      Cl.a27[(size_t)st27(0, 0, 0) * NC + I] = 1.0;
    }
  }
}

double diag_of(const Device& d, const MgHierarchy& h, int l, size_t n) {
  return l == 0 ? d.diag[n] : h.lv[l].a27[(size_t)st27(0, 0, 0) * h.lv[l].n + n];
}

// lambda_max of D^-1 A on the unknown nodes by power iteration from a fixed vector
double lambda_max(const Device& d, const MgHierarchy& h, int l) {
  const MgLevel& L = h.lv[l];
  std::vector<double> y(L.n), z(L.n);
  for (size_t n = 0; n < L.n; ++n) y[n] = L.fixed[n] ? 0.0 : 1.0 + 0.5 * std::sin(0.7 * (double)n);
  double lam = 0;
  for (int it = 0; it < 40; ++it) {
    mg_apply(d, h, l, y.data(), z.data());
    double nz2 = 0, yz = 0, yy = 0;
    for (size_t n = 0; n < L.n; ++n) {
      if (L.fixed[n]) { z[n] = 0; continue; }
      z[n] /= diag_of(d, h, l, n);
      yz += y[n] * z[n];
      yy += y[n] * y[n];
      nz2 += z[n] * z[n];
    }
    lam = yy > 0 ? yz / yy : 0;
    const double s = nz2 > 0 ? 1.0 / std::sqrt(nz2) : 0.0;
    for (size_t n = 0; n < L.n; ++n) y[n] = z[n] * s;
  }
  return lam;
}

// dense inverse of an SPD matrix by Cholesky (n small)
bool spd_inverse(std::vector<double>& a, int n, int threads) {
  // Cholesky A = L L^T, row by row (rows of L contiguous)
  std::vector<double> L((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    const double* lj = &L[(size_t)j * n];
    double s = a[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= lj[k] * lj[k];
    if (!(s > 0)) return false;
    const double ljj = std::sqrt(s);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      const double* li = &L[(size_t)i * n];
      double t = a[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= li[k] * lj[k];
      L[(size_t)i * n + j] = t / ljj;
    }
  }
  // U = L^T stored by rows, so the back substitution also reads rows
  std::vector<double> U((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k <= i; ++k) U[(size_t)k * n + i] = L[(size_t)i * n + k];
  // column c of the inverse: L y = e_c (y_i = 0 for i < c), then L^T x = y;
  // columns are independent and split over the threads
  std::vector<double> inv((size_t)n * n);
  auto work = [&](int c0, int c1) {
    std::vector<double> y(n), x(n);
    for (int c = c0; c < c1; ++c) {
      for (int i = 0; i < c; ++i) y[i] = 0.0;
      for (int i = c; i < n; ++i) {
        const double* li = &L[(size_t)i * n];
        double t = i == c ? 1.0 : 0.0;
        for (int k = c; k < i; ++k) t -= li[k] * y[k];
        y[i] = t / li[i];
      }
      for (int i = n - 1; i >= 0; --i) {
        const double* ui = &U[(size_t)i * n];
        double t = y[i];
        for (int k = i + 1; k < n; ++k) t -= ui[k] * x[k];
        x[i] = t / ui[i];
      }
      for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = x[i];
    }
  };
  const int nth = std::max(1, std::min(threads, n / 64 + 1));
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t) th.emplace_back(work, (int)((long long)n * t / nth), (int)((long long)n * (t + 1) / nth));
  for (auto& t : th) t.join();
  a.swap(inv);
  return true;
}

}  // namespace

void mg_apply(const Device& d, const MgHierarchy& h, int l, const double* x, double* y) {
  const MgLevel& L = h.lv[l];
  if (l == 0) {
    const size_t sx = L.nz, sy = (size_t)L.nx * L.nz;
    for (int j = 0; j < L.ny; ++j)
      for (int i = 0; i < L.nx; ++i)
        for (int k = 0; k < L.nz; ++k) {
          const size_t n = ((size_t)j * L.nx + i) * L.nz + k;
          double v = d.diag[n] * x[n];
          if (i + 1 < L.nx) v += d.cx[n] * x[n + sx];
          if (i > 0) v += d.cx[n - sx] * x[n - sx];
          if (j + 1 < L.ny) v += d.cy[n] * x[n + sy];
          if (j > 0) v += d.cy[n - sy] * x[n - sy];
          if (k + 1 < L.nz) v += d.cz[n] * x[n + 1];
          if (k > 0) v += d.cz[n - 1] * x[n - 1];
          y[n] = v;
        }
    return;
  }
  const Op op{d, L, false};
  for (int j = 0; j < L.ny; ++j)
    for (int i = 0; i < L.nx; ++i)
      for (int k = 0; k < L.nz; ++k) {
        const size_t n = ((size_t)j * L.nx + i) * L.nz + k;
        double v = 0;
        for (int dj = -1; dj <= 1; ++dj)
          for (int di = -1; di <= 1; ++di)
            for (int dk = -1; dk <= 1; ++dk) {
              const int gi = i + di, gj = j + dj, gk = k + dk;
              if (gi < 0 || gi >= L.nx || gj < 0 || gj >= L.ny || gk < 0 || gk >= L.nz) continue;
              const double a = op.at(n, i, j, k, di, dj, dk);
              if (a != 0.0) v += a * x[((size_t)gj * L.nx + gi) * L.nz + gk];
            }
        y[n] = v;
      }
}

std::string mg_build(const Device& d, MgHierarchy& h, int min_nodes, int threads) {
  h.lv.clear();
  h.lv.emplace_back();
  MgLevel& L0 = h.lv[0];
  L0.nx = d.pnx; L0.ny = d.pny; L0.nz = d.pnz; L0.n = d.nnode;
  L0.fixed.resize(L0.n);
  for (size_t n = 0; n < L0.n; ++n) L0.fixed[n] = d.node_kind[n] != 0;
  while (true) {
    const MgLevel& F = h.lv.back();
    if ((long long)F.n <= min_nodes || std::min(F.nx, std::min(F.ny, F.nz)) < 3) break;
    MgLevel C;
    C.nx = F.nx / 2 + 1; C.ny = F.ny / 2 + 1; C.nz = F.nz / 2 + 1;
    C.n = (size_t)C.nx * C.ny * C.nz;
    galerkin(d, F, h.lv.size() == 1, C, threads);
    h.lv.push_back(std::move(C));
  }
  for (size_t l = 0; l < h.lv.size(); ++l) {
    MgLevel& L = h.lv[l];
    L.lambda_max = lambda_max(d, h, (int)l);
    L.omega = L.lambda_max > 0 ? 4.0 / (3.0 * L.lambda_max) : 0.0;
    L.wdinv.assign(L.n, 0.0);
    for (size_t n = 0; n < L.n; ++n)
      if (!L.fixed[n]) L.wdinv[n] = L.omega / diag_of(d, h, (int)l, n);
  }
  // dense inverse of the coarsest operator (fixed rows are identity rows)
  const int lc = (int)h.lv.size() - 1;
  const MgLevel& Lc = h.lv[lc];
  const Op op{d, Lc, lc == 0};
  const int nc = (int)Lc.n;
  std::vector<double> a((size_t)nc * nc, 0.0);
  for (int j = 0; j < Lc.ny; ++j)
    for (int i = 0; i < Lc.nx; ++i)
      for (int k = 0; k < Lc.nz; ++k) {
        const size_t n = ((size_t)j * Lc.nx + i) * Lc.nz + k;
        for (int dj = -1; dj <= 1; ++dj)
          for (int di = -1; di <= 1; ++di)
            for (int dk = -1; dk <= 1; ++dk) {
              const int gi = i + di, gj = j + dj, gk = k + dk;
              if (gi < 0 || gi >= Lc.nx || gj < 0 || gj >= Lc.ny || gk < 0 || gk >= Lc.nz) continue;
              a[n * nc + ((size_t)gj * Lc.nx + gi) * Lc.nz + gk] = op.at(n, i, j, k, di, dj, dk);
            }
      }
  // The Galerkin product is positive semi-definite; where the interpolation of
  // two coarse nodes sees the same unknown fine nodes only (ghost nodes past an
  // even-sized edge next to fixed planes) it is singular in exact arithmetic, and
  // the factorisation meets a zero pivot.  The coarsest solve only preconditions:
  // it is then done on A_c + delta I with the smallest delta (1e-12, 1e-9, 1e-6 of
  // the largest diagonal entry) that factorises, which keeps it SPD.
  double dmax = 0.0;
  for (int q = 0; q < nc; ++q) dmax = std::max(dmax, std::fabs(a[(size_t)q * nc + q]));
  std::vector<double> a0 = a;
  h.coarse_shift = 0.0;
  bool ok = spd_inverse(a, nc, threads);
  for (double f : {1e-12, 1e-9, 1e-6}) {
    if (ok) break;
    a = a0;
    for (int q = 0; q < nc; ++q) a[(size_t)q * nc + q] += f * dmax;
    h.coarse_shift = f * dmax;
    ok = spd_inverse(a, nc, threads);
  }
  if (!ok) return "coarsest multigrid operator is not positive definite";
  h.cinv.swap(a);
  return "";
}

}  // namespace emc
