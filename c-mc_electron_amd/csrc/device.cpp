// Deck loading and mesh arrays (see device.h for the reference map).
#include "device.h"

#include <cmath>
#include <fstream>
#include <sstream>

namespace emc {

static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

static std::string dirname_of(const std::string& path) {
  size_t p = path.find_last_of('/');
  if (p == std::string::npos) return ".";
  return path.substr(0, p);
}

static std::string join(const std::string& dir, const std::string& f) {
  if (!f.empty() && f[0] == '/') return f;
  return dir + "/" + f;
}

// input.txt: "key = value" lines, '#' starts a comment (finfet/input.txt:49-54).
static std::string parse_input(Config& c, const std::string& path) {
  std::ifstream f(path);
  if (!f) return "cannot open " + path;
  std::string line;
  while (std::getline(f, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    size_t e = line.find('=');
    if (e == std::string::npos) continue;
    std::string k = trim(line.substr(0, e)), v = trim(line.substr(e + 1));
    if (!k.empty()) c.raw[k] = v;
  }
  auto gi = [&](const char* k, int& out) {
    auto it = c.raw.find(k);
    if (it != c.raw.end()) out = (int)std::stod(it->second);
  };
  auto gd = [&](const char* k, double& out) {
    auto it = c.raw.find(k);
    if (it != c.raw.end()) out = std::stod(it->second);
  };
  auto gb = [&](const char* k, bool& out) {
    auto it = c.raw.find(k);
    if (it == c.raw.end()) return;
    const std::string& v = it->second;
    if (v == "true" || v == "True") out = true;
    else if (v == "false" || v == "False") out = false;
    else out = std::stod(v) != 0.0;
  };
  auto gs = [&](const char* k, std::string& out) {
    auto it = c.raw.find(k);
    if (it != c.raw.end()) out = it->second;
  };
  try {
    gs("gridFile", c.grid_file);
    gs("device_file_name", c.device_file);
    gi("total_step", c.total_step);
    gi("ElectronNumber", c.electron_number);
    gi("HoleNumber", c.hole_number);
    gd("Temperature", c.temperature);
    gd("dt", c.dt);
    gi("stat_step", c.stat_step);
    gi("stat_heat_step", c.stat_heat_step);
    gi("heat_steps", c.heat_steps);
    gi("mr_step", c.mr_step);
    gb("Flag_MultipleRefresh", c.flag_mr);
    gb("Flag_SurfaceScatter", c.flag_surface);
    gb("Flag_SurfaceRoughnessScatter", c.flag_rough);
    gb("Flag_SurfacePhononScatter", c.flag_sphonon);
    gb("Flag_SurfaceImpurityScatter", c.flag_simp);
    gb("Flag_compute_heat", c.flag_heat);
    gb("Flag_QuantumCorrection", c.flag_qc);
    gb("Flag_restart", c.flag_restart);
    gi("restart_step", c.restart_step);
    gs("restart_filename", c.restart_file);
    gi("quantum_start_step", c.quantum_start_step);
  } catch (...) {
    return "input.txt: malformed numeric value";
  }
  if (c.stat_step <= 0) return "input.txt: stat_step must be positive";
  return "";
}

// lgrid.txt: counts and coordinates [um] per axis (src/mcmodel.cpp:4595-4693).
static std::string parse_grid(Device& d, const std::string& path) {
  std::ifstream f(path);
  if (!f) {
    // the reference also tries finfet/<name> (src/mcmodel.cpp:4610-4614)
    f.open(join(dirname_of(path), "finfet/" + path.substr(path.find_last_of('/') + 1)));
    if (!f) return "cannot open grid file " + path;
  }
  auto axis = [&](int& n, std::vector<double>& l, std::vector<double>& dl) -> bool {
    if (!(f >> n) || n < 2) return false;
    l.resize(n);
    for (int i = 0; i < n; ++i) {
      double pos;
      if (!(f >> pos)) return false;
      l[i] = pos * 1e-6;
    }
    dl.resize(n - 1);
    for (int i = 1; i < n; ++i) dl[i - 1] = l[i] - l[i - 1];
    return true;
  };
  if (!axis(d.pnx, d.lx, d.dx) || !axis(d.pny, d.ly, d.dy) || !axis(d.pnz, d.lz, d.dz))
    return "grid file " + path + ": malformed";
  d.cnx = d.pnx - 1; d.cny = d.pny - 1; d.cnz = d.pnz - 1;
  if (d.cnx > PK_MAXI + 1 || d.cny > PK_MAXJ + 1 || d.cnz > PK_MAXK + 1)
    return "grid too large for the packed cell index";
  for (double v : d.dx) if (!(v > 0)) return "grid x coordinates not increasing";
  for (double v : d.dy) if (!(v > 0)) return "grid y coordinates not increasing";
  for (double v : d.dz) if (!(v > 0)) return "grid z coordinates not increasing";
  d.ncell = (size_t)d.cnx * d.cny * d.cnz;
  d.nnode = (size_t)d.pnx * d.pny * d.pnz;
  return "";
}

// distance2index: deck positions are nm and must sit on a grid line within
// 1e-4 nm (src/mcmodel.cpp:4160-4177).
static int pos_index(double pos_nm, const std::vector<double>& cut_m) {
  for (size_t i = 0; i < cut_m.size(); ++i)
    if (std::fabs(pos_nm - cut_m[i] * 1e9) < 1e-4) return (int)i;
  return -1;
}

namespace {
struct Cmd {
  int type = 0;
  int range[6] = {0, 0, 0, 0, 0, 0};
  int ip[6] = {0, 0, 0, 0, 0, 0};
  double dp = 0;
};
}  // namespace

// init_vector_entry (src/set_user_defined_data.cpp:3-26): half-open cell ranges,
// all three extents must be non-empty.
template <class T>
static void fill_cells(Device& d, const int* r, T v, std::vector<T>& a) {
  if (r[0] >= 0 && r[1] <= d.pnx - 1 && r[0] < r[1] && r[4] >= 0 && r[5] <= d.pnz - 1 && r[4] < r[5] &&
      r[2] >= 0 && r[3] <= d.pny - 1 && r[2] < r[3]) {
    for (int i = r[0]; i < r[1]; ++i)
      for (int k = r[4]; k < r[5]; ++k)
        for (int j = r[2]; j < r[3]; ++j) a[d.cell(i, j, k)] = v;
  }
}

// SetMRCube (src/set_user_defined_data.cpp:52-80)
static void set_mr_cube(Device& d, const int* r, const int* mr) {
  if (r[0] >= 0 && r[1] <= d.pnx - 1 && r[0] < r[1] && r[4] >= 0 && r[5] <= d.pnz - 1 && r[4] < r[5] &&
      r[2] >= 0 && r[3] <= d.pny - 1 && r[2] < r[3]) {
    for (int i = r[0]; i < r[1]; ++i)
      for (int k = r[4]; k < r[5]; ++k)
        for (int j = r[2]; j < r[3]; ++j)
          for (int q = 0; q < 6; ++q) d.motion[d.cell(i, j, k) * 6 + q] = (uint8_t)mr[q];
  }
}

// SetMRPlane (src/set_user_defined_data.cpp:82-176): the rule applies to the
// cell on the side the particle comes from.
static void set_mr_plane(Device& d, const int* r, int dir, int mr) {
  const int bi = r[0], ei = r[1], bj = r[2], ej = r[3], bk = r[4], ek = r[5];
  if (!(bi >= 0 && ei <= d.pnx - 1 && bi <= ei && bk >= 0 && ek <= d.pnz - 1 && bk <= ek && bj >= 0 &&
        ej <= d.pny - 1 && bj <= ej))
    return;
  auto set = [&](int i, int j, int k) { d.motion[d.cell(i, j, k) * 6 + dir] = (uint8_t)mr; };
  if (bi == ei) {
    if (bj == ej || bk == ek) return;  // "Must be a Plane"
    if (dir == UP) {
      if (bi <= d.cnx - 1)
        for (int k = bk; k < ek; ++k)
          for (int j = bj; j < ej; ++j) set(bi, j, k);
    } else if (dir == DOWN) {
      if (bi > 0)
        for (int k = bk; k < ek; ++k)
          for (int j = bj; j < ej; ++j) set(bi - 1, j, k);
    }
  } else if (bj == ej) {
    if (bk == ek) return;
    if (dir == LEFT) {
      if (bj >= 0 && bj <= d.cny - 1)
        for (int k = bk; k < ek; ++k)
          for (int i = bi; i < ei; ++i) set(i, bj, k);
    } else if (dir == RIGHT) {
      if (bj - 1 >= 0 && bj - 1 <= d.cny - 1)
        for (int k = bk; k < ek; ++k)
          for (int i = bi; i < ei; ++i) set(i, bj - 1, k);
    }
  } else if (bk == ek) {
    if (dir == BACK) {
      if (bk > 0)
        for (int i = bi; i < ei; ++i)
          for (int j = bj; j < ej; ++j) set(i, j, bk - 1);
    } else if (dir == FRONT) {
      if (bk < d.cnz - 1)  // as in the reference: a FRONT rule on the last plane is ignored
        for (int i = bi; i < ei; ++i)
          for (int j = bj; j < ej; ++j) set(i, j, bk);
    }
  }
}

static std::string parse_ldg(Device& d, const std::string& path, std::vector<Cmd>& cmds) {
  std::ifstream f(path);
  if (!f) return "cannot open device file " + path;
  static const std::map<std::string, int> op = {
      {"donor", 1}, {"acceptor", 2}, {"region", 3}, {"motioncube", 4}, {"attachcontact", 5},
      {"contact", 6}, {"motionplane", 7}, {"parnumber", 8}, {"default_par_number", 9},
      {"ScatterArea", 12}, {"VsVdVg", 13}, {"vgrange", 14}, {"ep_parm", 15}, {"surfaces", 16},
      {"surface_scatter_range", 17}, {"quantumRegion", 18}};
  static const std::map<std::string, int> dirs = {{"UP", UP}, {"DOWN", DOWN}, {"LEFT", LEFT},
                                                  {"RIGHT", RIGHT}, {"FRONT", FRONT}, {"BACK", BACK}};
  static const std::map<std::string, int> rules = {
      {"PASS", PASS}, {"REFLECT", REFLECT}, {"SCATTOX", SCATTOX}, {"GENERATE", GENERATE},
      {"CATCH", CATCH}, {"PERIOD", PERIOD}, {"GENREF", GENREF}, {"CATCHGATE", CATCHGATE}};
  static const std::map<std::string, int> mats = {
      {"VACUUM", VACUUM}, {"OXIDE", OXIDE}, {"SILICON", SILICON}, {"IGZO", IGZO}};
  auto lookup = [](const std::map<std::string, int>& m, const std::string& k) {
    auto it = m.find(k);
    return it == m.end() ? 0 : it->second;
  };
  std::string err;
  auto cube = [&](int* idx) -> bool {
    double pos[6];
    for (int q = 0; q < 6; ++q)
      if (!(f >> pos[q])) { err = "device file: truncated coordinates"; return false; }
    const std::vector<double>* ax[6] = {&d.lx, &d.lx, &d.ly, &d.ly, &d.lz, &d.lz};
    for (int q = 0; q < 6; ++q) {
      idx[q] = pos_index(pos[q], *ax[q]);
      if (idx[q] < 0) {
        std::ostringstream o;
        o << "device file: coordinate " << pos[q] << " not aligned with the grid";
        err = o.str();
        return false;
      }
    }
    return true;
  };
  std::string key;
  if (!(f >> key)) return "device file: empty";
  while (key != "end") {
    Cmd c;
    c.type = lookup(op, key);
    std::string s;
    switch (c.type) {
      case 1: case 2:
        if (!cube(c.range)) return err;
        f >> c.dp;
        break;
      case 3:
        if (!cube(c.range)) return err;
        f >> s;
        c.ip[0] = lookup(mats, s);
        if (c.ip[0] == SILICON) { c.ip[1] = d.default_e; c.ip[2] = d.default_h; }
        else if (c.ip[0] == IGZO) { c.ip[1] = d.default_e; c.ip[2] = d.default_e; }
        break;
      case 4:
        if (!cube(c.range)) return err;
        for (int q = 0; q < 6; ++q) { f >> s; c.ip[q] = lookup(rules, s); }
        break;
      case 5: case 12:
        if (!cube(c.range)) return err;
        f >> c.ip[0];
        break;
      case 6: {
        ContactDef ct;
        f >> ct.nplane >> ct.phims;
        if (ct.nplane < 0 || ct.nplane > MNContactPlane) return "device file: bad contact plane count";
        for (int q = 0; q < ct.nplane; ++q) {
          int idx[6];
          if (!cube(idx)) return err;
          ct.bi[q] = idx[0]; ct.ei[q] = idx[1]; ct.bj[q] = idx[2];
          ct.ej[q] = idx[3]; ct.bk[q] = idx[4]; ct.ek[q] = idx[5];
        }
        f >> ct.vapp;
        if ((int)d.contacts.size() >= MNContact) return "device file: too many contacts";
        d.contacts.push_back(ct);
        break;
      }
      case 7:
        if (!cube(c.range)) return err;
        f >> s; c.ip[0] = lookup(dirs, s);
        f >> s; c.ip[1] = lookup(rules, s);
        break;
      case 8:
        if (!cube(c.range)) return err;
        f >> c.ip[0] >> c.ip[1];
        break;
      case 9:
        f >> d.default_e >> d.default_h;
        break;
      case 13: case 14: {
        double t[4];
        f >> t[0] >> t[1] >> t[2] >> t[3];
        break;
      }
      case 15:  // ep_parm qc_xratio qc_xtheta qc_yratio qc_ytheta qc_zratio qc_ztheta Eb [nm, eV]
        f >> d.qc_ratio[0] >> d.qc_theta[0] >> d.qc_ratio[1] >> d.qc_theta[1] >> d.qc_ratio[2] >> d.qc_theta[2] >>
            d.qc_eb;
        d.has_ep = true;
        break;
      case 16:
        f >> d.nsurf;
        if (d.nsurf < 0 || d.nsurf > MNSurface) return "device file: bad surface count";
        for (int q = 0; q < d.nsurf; ++q) {
          double pos;
          f >> d.surf_type[q] >> pos >> d.surf_dir[q];
          d.surf_pos[q] = pos * 1e-9;  // metres; normalised with the grid
        }
        break;
      case 17: case 18:
        if (!cube(c.range)) return err;
        break;
      default:
        break;  // unknown token: skipped, as the reference's op[] lookup yields 0
    }
    if (!f) return "device file: malformed command '" + key + "'";
    cmds.push_back(c);
    if (!(f >> key)) return "device file: missing 'end'";
  }
  return "";
}

std::string device_parse(Device& d, const std::string& input_path) {
  d.cfg = Config();
  d.cfg.dir = dirname_of(input_path);
  std::string err = parse_input(d.cfg, input_path);
  if (!err.empty()) return err;
  phys_init(d.phys, d.cfg.temperature);
  err = parse_grid(d, join(d.cfg.dir, d.cfg.grid_file));
  if (!err.empty()) return err;
  const size_t nc = d.ncell;
  d.material.assign(nc, 0); d.attached.assign(nc, 0); d.desired_e.assign(nc, 0);
  d.desired_h.assign(nc, 0); d.scatter_type.assign(nc, 0); d.motion.assign(nc * 6, 0);
  d.insurf.assign(nc, 0); d.qcregion.assign(nc, 0); d.donor.assign(nc, 0.0); d.acceptor.assign(nc, 0.0);
  d.contacts.clear();
  std::vector<Cmd> cmds;
  err = parse_ldg(d, join(d.cfg.dir, d.cfg.device_file), cmds);
  if (!err.empty()) return err;
  // init_by_user_input (src/mcmodel.cpp:3065-3153), in deck order
  for (const Cmd& c : cmds) {
    switch (c.type) {
      case 1: fill_cells(d, c.range, c.dp, d.donor); break;
      case 2: fill_cells(d, c.range, c.dp, d.acceptor); break;
      case 3:
        fill_cells(d, c.range, c.ip[0], d.material);
        fill_cells(d, c.range, c.ip[1], d.desired_e);
        fill_cells(d, c.range, c.ip[2], d.desired_h);
        break;
      case 4: set_mr_cube(d, c.range, c.ip); break;
      case 5: fill_cells(d, c.range, c.ip[0], d.attached); break;
      case 7: set_mr_plane(d, c.range, c.ip[0], c.ip[1]); break;
      case 8:
        fill_cells(d, c.range, c.ip[0], d.desired_e);
        fill_cells(d, c.range, c.ip[1], d.desired_h);
        break;
      case 12: fill_cells(d, c.range, c.ip[0], d.scatter_type); break;
      case 17:  // surface_scatter_range falls through into quantumRegion (src/mcmodel.cpp:3140-3149)
        fill_cells(d, c.range, (uint8_t)1, d.insurf);
        fill_cells(d, c.range, (uint8_t)1, d.qcregion);
        break;
      case 18: fill_cells(d, c.range, (uint8_t)1, d.qcregion); break;
      default: break;
    }
  }
  // scaling (src/mcmodel.cpp:4125-4158)
  const Phys& p = d.phys;
  for (auto* v : {&d.dx, &d.dy, &d.dz, &d.lx, &d.ly, &d.lz})
    for (double& x : *v) x /= p.spr0;
  d.cfg.dt /= p.time0;
  for (auto& ct : d.contacts) { ct.vapp /= p.pot0; ct.phims /= p.pot0; }
  for (size_t c = 0; c < nc; ++c) { d.donor[c] /= p.conc0; d.acceptor[c] /= p.conc0; }
  for (int q = 0; q < d.nsurf; ++q) d.surf_pos[q] /= p.spr0;
  for (double& t : d.qc_theta) t /= 1e9 * p.spr0;  // src/mcmodel.cpp:4550-4555
  d.qc_eb /= p.pot0;
  if (d.cfg.flag_qc && !d.has_ep) return "device file: Flag_QuantumCorrection needs ep_parm";
  if (d.cfg.flag_qc && !(d.qc_theta[0] > 0 && d.qc_theta[1] > 0 && d.qc_theta[2] > 0))
    return "device file: ep_parm lengths must be positive";
  return "";
}

static void compute_dirichlet(Device& d);

std::string device_finish(Device& d, const Band& band) {
  const Phys& p = d.phys;
  const size_t nc = d.ncell, nn = d.nnode;
  // init_cell_data (src/mcmodel.cpp:3187-3268)
  d.cvol.assign(nc, 0.0); d.da.assign(nc, 0.0);
  d.init_e_charge.assign(nc, 0.0); d.init_h_charge.assign(nc, 0.0);
  d.sum_charge[0] = d.sum_charge[1] = 0.0;
  for (int i = 0; i < d.cnx; ++i)
    for (int k = 0; k < d.cnz; ++k)
      for (int j = 0; j < d.cny; ++j) {
        size_t c = d.cell(i, j, k);
        d.cvol[c] = d.dx[i] * d.dy[j] * d.dz[k];
        d.da[c] = d.donor[c] + d.acceptor[c];
      }
  const double Ni = band.Ni;
  for (int i = 0; i < d.cnx; ++i)
    for (int k = 0; k < d.cnz; ++k)
      for (int j = 0; j < d.cny; ++j) {
        size_t c = d.cell(i, j, k);
        if (d.material[c] != SILICON) continue;
        double dope = d.donor[c] - d.acceptor[c];
        if (dope > 0) {
          d.init_e_charge[c] = -dope * d.cvol[c];
          d.init_h_charge[c] = (Ni * Ni / dope) * d.cvol[c];
        } else if (dope < 0) {
          d.init_e_charge[c] = (Ni * Ni / dope) * d.cvol[c];
          d.init_h_charge[c] = -dope * d.cvol[c];
        } else {
          d.init_e_charge[c] = -Ni * d.cvol[c];
          d.init_h_charge[c] = Ni * d.cvol[c];
        }
        d.sum_charge[0] += d.init_e_charge[c];
        d.sum_charge[1] += d.init_h_charge[c];
      }
  // init_surface_roughness (src/mcmodel.cpp:3155-3185); note the reference
  // measures from the cell's lower grid line, not its centre.
  d.nearest_surf.assign(nc, 0); d.eeff_dir.assign(nc, 0);
  if (d.cfg.flag_surface) {
    for (int i = 0; i < d.cnx; ++i)
      for (int k = 0; k < d.cnz; ++k)
        for (int j = 0; j < d.cny; ++j) {
          size_t c = d.cell(i, j, k);
          double dist = 1e99;
          for (int s = 0; s < d.nsurf; ++s) {
            double q = d.surf_type[s] == 0 ? d.lx[i] : d.surf_type[s] == 1 ? d.ly[j] : d.lz[k];
            if ((d.surf_type[s] == 0 || d.surf_type[s] == 1 || d.surf_type[s] == 2) &&
                dist > std::fabs(q - d.surf_pos[s])) {
              dist = std::fabs(q - d.surf_pos[s]);
              d.nearest_surf[c] = s;
            }
          }
          d.eeff_dir[c] = d.surf_type[d.nearest_surf[c]];
        }
  }
  // init_point_data (src/mcmodel.cpp:3272-3390)
  d.pvol.assign(nn, 0.0); d.pdop.assign(nn, 0.0); d.vadd.assign(nn, 0.0);
  for (int i = 0; i < d.cnx; ++i)
    for (int k = 0; k < d.cnz; ++k)
      for (int j = 0; j < d.cny; ++j) {
        size_t c = d.cell(i, j, k);
        if (d.material[c] != SILICON) continue;
        double vt = d.cvol[c] * 0.125;
        double dat = vt * (d.donor[c] - d.acceptor[c]);
        d.pdop[d.node(i + 1, j, k)] += dat; d.pdop[d.node(i, j, k)] += dat;
        d.pdop[d.node(i + 1, j, k + 1)] += dat; d.pdop[d.node(i, j, k + 1)] += dat;
        d.pdop[d.node(i + 1, j + 1, k)] += dat; d.pdop[d.node(i, j + 1, k)] += dat;
        d.pdop[d.node(i + 1, j + 1, k + 1)] += dat; d.pdop[d.node(i, j + 1, k + 1)] += dat;
        d.pvol[d.node(i + 1, j, k)] += vt; d.pvol[d.node(i, j, k)] += vt;
        d.pvol[d.node(i + 1, j, k + 1)] += vt; d.pvol[d.node(i, j, k + 1)] += vt;
        d.pvol[d.node(i + 1, j + 1, k)] += vt; d.pvol[d.node(i, j + 1, k)] += vt;
        d.pvol[d.node(i + 1, j + 1, k + 1)] += vt; d.pvol[d.node(i, j + 1, k + 1)] += vt;
      }
  for (int i = 0; i < d.pnx; ++i)
    for (int k = 0; k < d.pnz; ++k)
      for (int j = 0; j < d.pny; ++j) {
        size_t n = d.node(i, j, k);
        double ndop = std::fabs(d.pvol[n]) > 1e-16 ? d.pdop[n] / d.pvol[n] : 0;
        double alfa = ndop / (Ni * 2.0);
        if (alfa > 0.0) d.vadd[n] = std::log(alfa + std::sqrt(1.0 + alfa * alfa));
        else if (alfa < 0) d.vadd[n] = -std::log(std::sqrt(1.0 + alfa * alfa) - alfa);
        else d.vadd[n] = 0;
      }
  for (int i = 0; i < d.pnx; ++i)
    for (int k = 0; k < d.pnz; ++k) {
      d.vadd[d.node(i, d.pny - 1, k)] = d.vadd[d.node(i, d.pny - 2, k)];
      d.vadd[d.node(i, 0, k)] = d.vadd[d.node(i, 1, k)];
    }
  // init_matrix (src/mcmodel.cpp:4695-4869): per-cell FV contributions to its
  // eight corners, cell loop order i, k, j.
  d.stencil.assign(nn * 7, 0.0);
  for (int i = 0; i < d.cnx; ++i)
    for (int k = 0; k < d.cnz; ++k)
      for (int j = 0; j < d.cny; ++j) {
        size_t c = d.cell(i, j, k);
        if (d.material[c] == VACUUM) continue;
        double e0 = p.eps[d.material[c]];
        double AA = -0.25 * e0 * d.dx[i] * d.dy[j] / d.dz[k];
        double BB = -0.25 * e0 * d.dx[i] * d.dz[k] / d.dy[j];
        double CC = -0.25 * e0 * d.dz[k] * d.dy[j] / d.dx[i];
        double DD = -AA - BB - CC;
        auto add = [&](int ii, int jj, int kk, int ex, int ey, int ez) {
          double* s = &d.stencil[d.node(ii, jj, kk) * 7];
          s[ez] += AA; s[ey] += BB; s[ex] += CC; s[6] += DD;
        };
        add(i, j, k, 1, 3, 5);
        add(i, j, k + 1, 1, 3, 4);
        add(i + 1, j, k, 0, 3, 5);
        add(i + 1, j, k + 1, 0, 3, 4);
        add(i, j + 1, k, 1, 2, 5);
        add(i, j + 1, k + 1, 1, 2, 4);
        add(i + 1, j + 1, k, 0, 2, 5);
        add(i + 1, j + 1, k + 1, 0, 2, 4);
      }
  d.node_kind.assign(nn, 0);
  for (size_t n = 0; n < nn; ++n) {
    bool any = false;
    for (int e = 0; e < 7; ++e) any = any || d.stencil[n * 7 + e] != 0;
    if (!any) { d.stencil[n * 7 + 6] = 1; d.node_kind[n] = 2; }
  }
  for (const ContactDef& ct : d.contacts)
    for (int q = 0; q < ct.nplane; ++q)
      for (int i = ct.bi[q]; i <= ct.ei[q]; ++i)
        for (int k = ct.bk[q]; k <= ct.ek[q]; ++k)
          for (int j = ct.bj[q]; j <= ct.ej[q]; ++j) {
            size_t n = d.node(i, j, k);
            for (int e = 0; e < 7; ++e) d.stencil[n * 7 + e] = 0;
            d.stencil[n * 7 + 6] = 1;
            d.node_kind[n] = 1;
          }
  // symmetric elimination of contact nodes -> SPD operator on unknown nodes
  d.diag.assign(nn, 1.0); d.cx.assign(nn, 0.0); d.cy.assign(nn, 0.0); d.cz.assign(nn, 0.0);
  for (int j = 0; j < d.pny; ++j)
    for (int i = 0; i < d.pnx; ++i)
      for (int k = 0; k < d.pnz; ++k) {
        size_t n = d.node(i, j, k);
        if (d.node_kind[n] != 0) continue;
        d.diag[n] = d.stencil[n * 7 + 6];
        if (i + 1 < d.pnx && d.node_kind[d.node(i + 1, j, k)] == 0) d.cx[n] = d.stencil[n * 7 + 1];
        if (j + 1 < d.pny && d.node_kind[d.node(i, j + 1, k)] == 0) d.cy[n] = d.stencil[n * 7 + 3];
        if (k + 1 < d.pnz && d.node_kind[d.node(i, j, k + 1)] == 0) d.cz[n] = d.stencil[n * 7 + 5];
      }
  compute_dirichlet(d);
  return "";
}

void device_set_bias(Device& d, const std::vector<double>& vapp_volts) {
  for (size_t q = 0; q < d.contacts.size() && q < vapp_volts.size(); ++q)
    d.contacts[q].vapp = vapp_volts[q] / d.phys.pot0;
  compute_dirichlet(d);
}

static void compute_dirichlet(Device& d) {
  const size_t nn = d.nnode;
  // compute_rhs contact rows (src/mcmodel.cpp:2883-2892): later contacts overwrite
  d.dirichlet_val.assign(nn, 0.0);
  for (const ContactDef& ct : d.contacts)
    for (int q = 0; q < ct.nplane; ++q)
      for (int i = ct.bi[q]; i <= ct.ei[q]; ++i)
        for (int k = ct.bk[q]; k <= ct.ek[q]; ++k)
          for (int j = ct.bj[q]; j <= ct.ej[q]; ++j) {
            size_t n = d.node(i, j, k);
            d.dirichlet_val[n] = ct.vapp + d.vadd[n] + ct.phims;
          }
  d.rhs_corr.assign(nn, 0.0);
  for (int j = 0; j < d.pny; ++j)
    for (int i = 0; i < d.pnx; ++i)
      for (int k = 0; k < d.pnz; ++k) {
        size_t n = d.node(i, j, k);
        if (d.node_kind[n] != 0) continue;
        const double* s = &d.stencil[n * 7];
        double corr = 0.0;
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (int e = 0; e < 6; ++e) {
          int ii = nb[e][0], jj = nb[e][1], kk = nb[e][2];
          if (ii < 0 || jj < 0 || kk < 0 || ii >= d.pnx || jj >= d.pny || kk >= d.pnz) continue;
          size_t m = d.node(ii, jj, kk);
          if (d.node_kind[m] == 1) corr += s[e] * d.dirichlet_val[m];
        }
        d.rhs_corr[n] = corr;
      }
}

std::string device_load(Device& d, const std::string& input_path, const Band* band) {
  std::string err = device_parse(d, input_path);
  if (!err.empty() || !band) return err;
  return device_finish(d, *band);
}

}  // namespace emc
