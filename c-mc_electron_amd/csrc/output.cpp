// data/ writers (src/mcmodel.cpp:756-842, 895-1161; src/utils.cpp:74-80, 311-338).
//
// Numbers are printed as the reference's ofstream prints them: doubles in the
// default floatfield at precision 6 (the %g conversion), integers in decimal,
// fields separated by one space, lines ended by '\n'.  Node files run i, then
// k, then j fastest (print_p_data's loop order); cell files run i, j, k.
#include "output.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "emc_const.h"

namespace emc {
namespace {

// A line buffer written to the file in one piece.
struct Text {
  std::string s;
  void num(double v) {
    char t[40];
    const int k = std::snprintf(t, sizeof t, "%g", v);
    s.append(t, (size_t)k);
  }
  void num(long long v) {
    char t[24];
    const int k = std::snprintf(t, sizeof t, "%lld", v);
    s.append(t, (size_t)k);
  }
  void num(int v) { num((long long)v); }
  Text& operator<<(const char* t) { s += t; return *this; }
  Text& operator<<(char c) { s += c; return *this; }
  Text& operator<<(double v) { num(v); return *this; }
  Text& operator<<(long long v) { num(v); return *this; }
  Text& operator<<(int v) { num(v); return *this; }
};

// getFileName (src/utils.cpp:74-80): prefix followed by the decimal number.
std::string file_name(const std::string& dir, const char* prefix, int num) {
  return dir + "/" + prefix + std::to_string(num);
}

std::string put_file(const std::string& path, const Text& t, bool append) {
  FILE* f = std::fopen(path.c_str(), append ? "ab" : "wb");
  if (!f) return "cannot open " + path + " for writing";
  const bool ok = std::fwrite(t.s.data(), 1, t.s.size(), f) == t.s.size();
  if (std::fclose(f) != 0 || !ok) return "writing " + path + " failed";
  return "";
}

inline double fx(int64_t v, double scale) { return (double)v / scale; }

}  // namespace

std::string write_stat(const Device& d, const StatWindow& w, bool heat_file, int step, const std::string& dir,
                       SavedStat& saved) {
  const Phys& ph = d.phys;
  const size_t NN = d.nnode;
  const int stat_step = d.cfg.stat_step;
  const double dts = d.cfg.dt * stat_step;
  std::string e;

  // current_scatter_info (src/mcmodel.cpp:756-842).  EnergyCatch is never
  // tallied by the reference (CatchAtContact adds to EnergyGen), so it prints 0.
  if (w.tallies) {
    const emc_fly_tallies& t = *w.tallies;
    Text c;
    c << "step = " << step << '\n';
    for (size_t q = 0; q < d.contacts.size(); ++q) {
      const double gen = fx(t.contact[q][EMC_CT_QGEN], EMC_FX_CONTACT);
      const double cat = fx(t.contact[q][EMC_CT_QCATCH], EMC_FX_CONTACT);
      const double egen = fx(t.contact[q][EMC_CT_EGEN], EMC_FX_CONTACT);
      const double ecat = 0.0;
      c << "icont = " << (int)q << '\n';
      c << "CharGen = " << gen << ' ' << "NumParGen = " << (long long)t.contact[q][EMC_CT_NGEN] << ' '
        << "CharCatch = " << cat << ' ' << "NumParCatch = " << (long long)t.contact[q][EMC_CT_NCATCH] << '\n';
      c << "EnergyGen= " << egen << '\n';
      c << "EnergyCatch= " << ecat << '\n';
      c << "EnergyCurrent= " << (ecat - egen) / dts * ph.pot0 << '\n';
      c << "Current = " << (gen - cat) / dts * ph.curr0 * 1e-2 << '\n';
    }
    c << "times of phonon scatter : " << (long long)t.n_phonon << '\n';
    c << "times of impurity scatter : " << (long long)t.n_impurity << '\n';
    if (!(e = put_file(dir + "/current", c, step != 0)).empty()) return e;
  }
  if (w.rcur && w.lcur) {
    Text c;
    for (int i = 0; i < d.pny; ++i) {
      const double r = fx(w.rcur[i], EMC_FX_CONTACT), l = fx(w.lcur[i], EMC_FX_CONTACT);
      c << i << ' ' << r / dts * ph.curr0 << ' ' << l / dts * ph.curr0 << ' ' << (r - l) / dts * ph.curr0 << '\n';
    }
    if (!(e = put_file(file_name(dir, "section_current", step), c, false)).empty()) return e;
  }

  // output_stat (src/mcmodel.cpp:971-1161)
  const int64_t* S = w.node_stat;
  auto stat = [&](int q, size_t n) { return fx(S[(size_t)q * NN + n], EMC_FX_STAT); };
  if (saved.e_charge.size() != NN) saved.e_charge.assign(NN, 0.0);
  if (saved.h_charge.size() != NN) saved.h_charge.assign(NN, 0.0);
  if (heat_file && w.heat_e && w.heat_h && S) {
    Text c;
    for (int i = 0; i < d.pnx; ++i)
      for (int k = 0; k < d.pnz; ++k)
        for (int j = 0; j < d.pny; ++j) {
          const size_t n = d.node(i, j, k);
          const double qe = stat(3, n), qh = stat(8, n);
          const double v1 = std::fabs(qe) > MY_ZERO ? fx(w.heat_e[n], EMC_FX_HEAT) * ph.pot0 / qe : 0.0;
          const double v2 = std::fabs(qh) > MY_ZERO ? fx(w.heat_h[n], EMC_FX_HEAT) * ph.pot0 / qh : 0.0;
          c << i << ' ' << j << ' ' << k << ' ' << v1 << ' ' << v1 * saved.e_charge[n] << ' ' << v2 << ' '
            << v2 * saved.h_charge[n] << '\n';
        }
    if (!(e = put_file(file_name(dir, "heat", step), c, true)).empty()) return e;
  }
  if (w.cell_count) {
    Text c;
    for (int i = 0; i < d.cnx; ++i)
      for (int j = 0; j < d.cny; ++j)
        for (int k = 0; k < d.cnz; ++k) c << i << ' ' << j << ' ' << k << ' ' << (double)w.cell_count[d.cell(i, j, k)] / stat_step << '\n';
    if (!(e = put_file(file_name(dir, "electron_num", step), c, true)).empty()) return e;
  }
  if (w.stat_pot) {
    saved.pot.assign(NN, 0.0);
    Text c;
    for (int i = 0; i < d.pnx; ++i)
      for (int k = 0; k < d.pnz; ++k)
        for (int j = 0; j < d.pny; ++j) {
          const size_t n = d.node(i, j, k);
          saved.pot[n] = w.stat_pot[n] / stat_step;
          c << i << ' ' << j << ' ' << k << ' ' << saved.pot[n] * ph.pot0 << '\n';
        }
    if (!(e = put_file(file_name(dir, "pot", step), c, true)).empty()) return e;
  }
  // stat_qc_pot sums the potential after the quantum correction, which is the
  // Poisson potential itself while the correction is off
  const double* qsum = w.stat_qc_pot ? w.stat_qc_pot : w.stat_pot;
  if (qsum) {
    saved.qc_pot.assign(NN, 0.0);
    Text c;
    for (int i = 0; i < d.pnx; ++i)
      for (int k = 0; k < d.pnz; ++k)
        for (int j = 0; j < d.pny; ++j) {
          const size_t n = d.node(i, j, k);
          saved.qc_pot[n] = qsum[n] / stat_step;
          c << i << ' ' << j << ' ' << k << ' ' << saved.qc_pot[n] * ph.pot0 << '\n';
        }
    if (!(e = put_file(file_name(dir, "qc_pot", step), c, true)).empty()) return e;
  }
  if (S) {
    for (int t = 0; t < 2; ++t) {
      std::vector<double>& keep = t == 0 ? saved.e_charge : saved.h_charge;
      Text c;
      for (int i = 0; i < d.pnx; ++i)
        for (int k = 0; k < d.pnz; ++k)
          for (int j = 0; j < d.pny; ++j) {
            const size_t n = d.node(i, j, k);
            double charge = stat(5 * t + 3, n);
            const double a = std::fabs(charge);
            double vx = 0, vy = 0, vz = 0, en = 0;
            if (a > MY_ZERO) {
              vx = stat(5 * t + 0, n) / a * ph.velo0;
              vy = stat(5 * t + 1, n) / a * ph.velo0;
              vz = stat(5 * t + 2, n) / a * ph.velo0;
              en = stat(5 * t + 4, n) / a * ph.pot0;
            }
            const double vol = d.pvol[n];
            if (vol > MY_ZERO) charge = charge / vol / stat_step * ph.conc0; else charge = 0;
            keep[n] = charge;
            c << i << ' ' << j << ' ' << k << ' ' << vx << ' ' << vy << ' ' << vz << ' ' << en << ' ' << charge << '\n';
          }
      if (!(e = put_file(file_name(dir, t == 0 ? "Electron" : "Hole", step), c, true)).empty()) return e;
    }
  }
  return "";
}

std::string write_restart(const Device& d, const emc_particles& p, int64_t n, int step, const std::string& dir) {
  // per cell, |q| summed in ascending particle key (the order MultipleRefresh
  // keeps a cell's particles in), so the sums do not depend on slot order
  std::vector<int64_t> order;
  order.reserve((size_t)n);
  for (int64_t s = 0; s < n; ++s)
    if (p.i[s] >= 0) order.push_back(s);
  std::vector<size_t> cell(order.size());
  for (size_t q = 0; q < order.size(); ++q) cell[q] = d.cell(p.i[order[q]], p.j[order[q]], p.k[order[q]]);
  std::vector<size_t> idx(order.size());
  std::iota(idx.begin(), idx.end(), (size_t)0);
  std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    if (cell[a] != cell[b]) return cell[a] < cell[b];
    return p.key[order[a]] < p.key[order[b]];
  });
  const size_t NC = d.ncell;
  std::vector<double> ec(NC, 0.0), hc(NC, 0.0);
  std::vector<long long> en(NC, 0), hn(NC, 0);
  for (size_t q : idx) {
    const int64_t s = order[q];
    if (p.ptype[s] == 0) { ec[cell[q]] += std::fabs(p.charge[s]); ++en[cell[q]]; }
    else { hc[cell[q]] += std::fabs(p.charge[s]); ++hn[cell[q]]; }
  }
  Text c;
  for (int i = 0; i < d.cnx; ++i)
    for (int j = 0; j < d.cny; ++j)
      for (int k = 0; k < d.cnz; ++k) {
        const size_t q = d.cell(i, j, k);
        c << i << ' ' << j << ' ' << k << ' ' << ec[q] << ' ' << en[q] << ' ' << hc[q] << ' ' << hn[q] << '\n';
      }
  c << "-1 -1 -1\n";
  return put_file(file_name(dir, "par_info_restart", step), c, true);
}

std::string write_pvolume(const Device& d, const std::string& dir) {
  Text c;
  for (int i = 0; i < d.pnx; ++i)
    for (int k = 0; k < d.pnz; ++k)
      for (int j = 0; j < d.pny; ++j) c << i << ' ' << j << ' ' << k << ' ' << d.pvol[d.node(i, j, k)] * d.phys.conc0 << '\n';
  return put_file(dir + "/pvolume", c, true);
}

}  // namespace emc
