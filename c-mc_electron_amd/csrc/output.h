// The data/ writers of the reference, in its file formats: output_stat and
// current_scatter_info (src/mcmodel.cpp:756-842, 971-1161), output_for_restart
// (src/mcmodel.cpp:895-965) and print_p_data (src/utils.cpp:311-338).
//
// Host code only: the caller brings the window sums as host arrays (copied
// from the device once per window, or built by hand in a test).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/emc.h"
#include "device.h"

namespace emc {

// One statistics window as host arrays, in the fixed point of the device sums.
struct StatWindow {
  const double* stat_pot = nullptr;       // [nnode] sum over the window's steps of the potential
  const double* stat_qc_pot = nullptr;    // [nnode] the same for the corrected potential (null: stat_pot)
  const int64_t* node_stat = nullptr;     // [10][nnode] EMC_FX_STAT, layout of emc_deposit
  const int64_t* cell_count = nullptr;    // [ncell] sum of the per-cell particle counts
  const int64_t* heat_e = nullptr;        // [nnode] EMC_FX_HEAT
  const int64_t* heat_h = nullptr;
  const int64_t* rcur = nullptr;          // [pny] EMC_FX_CONTACT
  const int64_t* lcur = nullptr;
  const emc_fly_tallies* tallies = nullptr;  // contact columns and scatter counts
};

// What output_stat keeps from one window for the next one and for the heat
// phase: stat_pot_saved and stat_chargeE_saved / stat_chargeH_saved.
struct SavedStat {
  std::vector<double> pot;                // stat_pot / stat_step, normalised
  std::vector<double> qc_pot;             // stat_qc_pot_saved: stat_qc_pot / stat_step
  std::vector<double> e_charge, h_charge; // carrier density of the Electron / Hole files
};

// data/current (appended; truncated when step == 0), data/section_current<step>,
// data/heat<step> when heat_file, data/electron_num<step>, data/pot<step>,
// data/qc_pot<step>, data/Electron<step>, data/Hole<step>.  Updates `saved`.
// Returns an error message, empty on success.
std::string write_stat(const Device& d, const StatWindow& w, bool heat_file, int step, const std::string& dir,
                       SavedStat& saved);

// data/par_info_restart<step>: per cell |Qe| Ne |Qh| Nh of the population
// (slots with i < 0 are empty), then "-1 -1 -1".
std::string write_restart(const Device& d, const emc_particles& p, int64_t n, int step, const std::string& dir);

// data/pvolume: the node volumes times conc0 (src/mcmodel.cpp:513).
std::string write_pvolume(const Device& d, const std::string& dir);

}  // namespace emc
