// Host-side deck loading and device description.
//
// Restates the parts of MeshQuantities that turn input.txt / lgrid.txt / ldg.txt
// into mesh arrays (src/mcmodel.cpp:3824-3987 getInputData, 4190-4584
// read_device_file, 4595-4693 read_grid_file, 4125-4158 scaling,
// src/set_user_defined_data.cpp init_vector_entry/SetMRCube/SetMRPlane,
// src/mcmodel.cpp:3065-3390 init_by_user_input, init_surface_roughness,
// init_cell_data, init_point_data) and the Poisson operator of init_matrix
// (src/mcmodel.cpp:4695-4869).
//
// Index order is the reference's single-rank order: cell c = (j*cnx + i)*cnz + k,
// node n = (j*pnx + i)*pnz + k (C_LINDEX_GHOST_ONE / P_LINDEX_ONE, include/defs.h).
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "band.h"

namespace emc {

struct ContactDef {
  int nplane = 0;
  int bi[MNContactPlane], ei[MNContactPlane], bj[MNContactPlane], ej[MNContactPlane],
      bk[MNContactPlane], ek[MNContactPlane];
  double phims = 0, vapp = 0;  // normalised after scaling
};

struct Config {
  std::string dir;  // directory holding input.txt (decks are resolved relative to it)
  std::string grid_file = "lgrid.txt", device_file = "ldg.txt";
  int total_step = 100000, electron_number = 100000, hole_number = 100000;
  double temperature = 300.0, dt = 1e-16;  // dt normalised after scaling
  int stat_step = 5000, stat_heat_step = 5000, heat_steps = 20000, mr_step = 10;
  bool flag_mr = false, flag_surface = true, flag_rough = true, flag_sphonon = true;
  bool flag_simp = false, flag_heat = false, flag_qc = false, flag_restart = false;
  int restart_step = 0;
  int quantum_start_step = 20;
  std::string restart_file = "nosuchname";  // resolved against dir when relative
  std::map<std::string, std::string> raw;
};

struct Device {
  Phys phys;
  Config cfg;
  int cnx = 0, cny = 0, cnz = 0, pnx = 0, pny = 0, pnz = 0;
  size_t ncell = 0, nnode = 0;
  std::vector<double> lx, ly, lz, dx, dy, dz;  // normalised (metres / spr0)

  // cell arrays
  std::vector<int> material, attached, desired_e, desired_h, scatter_type;
  std::vector<uint8_t> motion;  // [ncell][6]
  std::vector<uint8_t> insurf, qcregion;
  std::vector<int> nearest_surf, eeff_dir;
  std::vector<double> donor, acceptor, da, cvol;
  std::vector<double> init_e_charge, init_h_charge;
  double sum_charge[2] = {0, 0};
  int default_e = 0, default_h = 0;

  // quantum correction: ep_parm (qc_*ratio, qc_*theta normalised, Eb normalised)
  bool has_ep = false;
  double qc_ratio[3] = {0, 0, 0}, qc_theta[3] = {0, 0, 0}, qc_eb = 0;

  int nsurf = 0;
  int surf_type[MNSurface] = {0}, surf_dir[MNSurface] = {0};
  double surf_pos[MNSurface] = {0};
  std::vector<ContactDef> contacts;

  // node arrays
  std::vector<double> pvol, pdop, vadd;

  // Poisson operator.  stencil[n*7 + e]: e = 0..5 neighbours (i-1,i+1,j-1,j+1,k-1,k+1),
  // 6 = diagonal, exactly as assembled by init_matrix.  After symmetric Dirichlet
  // elimination: node_kind 0 = unknown, 1 = contact (Dirichlet), 2 = identity row;
  // diag/cx/cy/cz hold the eliminated SPD operator (cx[n] couples n and n+x).
  std::vector<double> stencil;
  std::vector<uint8_t> node_kind;
  std::vector<double> dirichlet_val;   // contact potential per node (0 elsewhere)
  std::vector<double> diag, cx, cy, cz;
  std::vector<double> rhs_corr;        // sum over contact neighbours a_uv * phi_v

  size_t cell(int i, int j, int k) const { return ((size_t)j * cnx + i) * cnz + k; }
  size_t node(int i, int j, int k) const { return ((size_t)j * pnx + i) * pnz + k; }
};

// Load input.txt (and the decks it names) from directory `dir`, build every mesh
// array, the Poisson operator and the contact potentials.  `band` must already
// be built (Ni is needed for the equilibrium charge).  Returns "" on success.
std::string device_load(Device& d, const std::string& input_txt_path, const Band* band);

// Phase 1 of device_load: parse the decks and normalise (no band data needed).
std::string device_parse(Device& d, const std::string& input_txt_path);
// Phase 2: cell/node data and Poisson operator (needs band.Ni).
std::string device_finish(Device& d, const Band& band);

// Recompute the contact (Dirichlet) potentials and the RHS correction after a
// change of contact bias.  vapp_volts[i] is the applied voltage of contact i.
void device_set_bias(Device& d, const std::vector<double>& vapp_volts);

}  // namespace emc
