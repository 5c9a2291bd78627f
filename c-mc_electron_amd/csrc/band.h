// Host-side physics normalisation and full-band tables.
//
// Phys restates MeshQuantities::init_phpysical_parameter (src/mcmodel.cpp:4007-4122);
// Band restates the tables Band::IELEC builds (src/Band.cpp:4271-4476): READBS,
// VELOTET, HESSTET, SETMATSYM, SETNEIBSYM, CALCFAC, GETMAXAOVTET, BUILDLISTS,
// GETDOS, BUILDPHSCATT and GETINDENS.  Tables are stored flat, row-major, with the
// index order of the reference arrays so that the device repack and the oracle
// can address them the same way.
#pragma once
#include <string>
#include <vector>

#include "emc_const.h"

namespace emc {

struct Phys {
  double T0, Tn;
  double eV0, em0, hq0, ec0, rmom0, spr0, spk0, time0, velo0, cvr;
  double pot0, field0, conc0, dens0, dpc0, scrt0, curr0;
  double Nc, N_cur;
  double sia0, sirho, siul, siut, a0pi, sieg;
  double eps[MNMaterialType];
  double frickel, psi_si;
  double charge_sign[NumParType];
  double mass_eff[NumParType];
  // surface-scattering model parameters (src/mcmodel.cpp:4105-4121)
  double SurfSc_ail, SurfSc_delta, SurfSc_XIph, SurfSc_theta, SurfSc_gama;
  double SurfSc_Nbmod, SurfSc_Npf, SurfSc_Pft, SurfSc_Pfn, SurfSc_Kpha;
};

// Normalisation for lattice temperature T [K]; Si bulk parameters.
void phys_init(Phys& p, double T);

struct Band {
  // dispersive phonon tables (load_phonon_tables), empty unless loaded:
  // ptab[t][row * PTAB_COLS[t] + col] as the reference indexes them, ptab_t the
  // transpose (one energy bin's column contiguous) for the flight
  std::vector<double> ptab[PTAB_N], ptab_t[PTAB_N];
  // ---- sizes
  int nband[NumParType] = {0, 0, 0};
  int bandof[NumParType] = {0, 0, 0};
  int nbt = 0;
  int partyp[MNB] = {0};
  int nk = 0, nt = 0;

  // ---- k-space grid (normalised) and tetrahedra
  std::vector<double> xkk, ykk, zkk, eek;  // [nk]
  std::vector<int> tet;                    // [4][nt]  corner ik, energy ascending
  std::vector<int> neib;                   // [4][nt]  NeibrTet (>=0 tet, <0 wedge/BZ face code)
  std::vector<int> ibt;                    // [nt]     band index
  std::vector<double> xkct, ykct, zkct;    // [nt]     centres
  std::vector<double> vt;                  // [4*nt]   vx,vy,vz,e(corner0)
  std::vector<double> vgt;                 // [nt]     |v|
  std::vector<double> dnt;                 // [4][4][nt] (component x,y,z,d ; face ; tet)
  std::vector<double> faclow, facmedium, fachigh, maxaovtet, gamtet;  // [nt]

  // ---- 48 symmetry operations of the cubic BZ
  int matsym[6][NSYM];
  int indmat[729];     // [sx+1][sy+1][sz+1][p0][p1][p2] -> isym
  int newisym[5][NSYM];
  double xbz[5][NSYM], ybz[5][NSYM], zbz[5][NSYM];

  // ---- energy-binned tetrahedron lists: [bin][band] counts and starts
  double dlist = 0, dlists = 0;
  std::vector<int> ntlist, ptlist, tlist;     // MWLE bins
  std::vector<int> nslist, pslist, slist;     // MWLE bins, BZ-surface tets only
  std::vector<int> ntlists, ptlists, tlists;  // MWLES bins

  // ---- 1 meV energy tables
  double dtable = 0, emin = 0, emax = 0;
  std::vector<double> energy;     // [NTAB1]
  std::vector<double> dos;        // [MNB][NTAB1]
  std::vector<double> sumdos;     // [NTAB1][3]
  std::vector<double> dostetmax;  // [MNB][NTAB1]
  double DOSMAX[NumParType] = {0, 0, 0};

  // ---- scattering tables
  int scpre = 0, scprh = 0, scproe = 0;
  std::vector<double> scatte;    // [MNScaEle][NBE][NBE]   (process, final, initial)
  std::vector<double> dose;      // [MNScaEle][NBE][NTAB1]
  std::vector<double> scattiie;  // [NBE][NBE][NTAB1]      (final, initial, itab)
  std::vector<double> scatth;    // [MNScaHole][NBH][NBH]
  std::vector<double> dosh;      // [MNScaHole][NBH][NTAB1]
  std::vector<double> scattiih;  // [NBH][NBH][NTAB1]
  std::vector<double> sumscatt;  // [NTAB1][MNB]
  double gamma[NumParType] = {0, 0, 0};

  // ---- phonon / band parameters (src/Band.cpp:4338-4399)
  double temptag, templag, templog, temptaf, templaf, temptof, temphop, tempoeop;
  double dftag, dflag, dflog, dftaf, dflaf, dftof, dfelast, dfhop, dfhelast, dfoeop, dfoeelast;
  double iifacelec, iifachole, hiithresh, hiiexp, ephb, mell, melt, meld;
  double difpr[NumParType];
  double Ni = 0;  // intrinsic density (normalised)

  // ---- interpolators (src/Band.cpp:1999-2321), host copies used for table building
  double caldos(double e, int ib) const;
  double caldossum(double e, int ip) const;
  double calscattsum(double e, int ib) const;
  double surf(double e, int it) const;
  double fastsurf(double e, int it) const;
  void out_wedge(double& xo, double& yo, double& zo, double xi, double yi, double zi, int isym) const;
  void confbz(double& x, double& y, double& z) const;
  int tet_corner(int c, int it) const { return tet[(size_t)c * nt + it]; }
  double& dn(int comp, int face, int it) { return dnt[((size_t)comp * 4 + face) * nt + it]; }
  double dn(int comp, int face, int it) const { return dnt[((size_t)comp * 4 + face) * nt + it]; }

  double a0pi = 0, sieg = 0;
};

// Build every table from <input_dir>/bs.si.asc and <input_dir>/zd.si.asc.
// Returns an empty string on success, otherwise an error message.
std::string band_build(Band& b, const Phys& p, const std::string& input_dir);

// init_deep (src/ParticleUtils.cpp:750-808): the six dispersive phonon tables from
// `files` (kloem kloab ktoem ktoab klaem klaab) in input_dir, each stored column
// by column in the file.
std::string load_phonon_tables(Band& b, const std::string& input_dir, const std::string files[PTAB_N]);

// sign() of the reference (src/utils.cpp:64): +-1; the build defines sign(0)=+1.
inline int sgn(double x) { return x < 0 ? -1 : 1; }

}  // namespace emc
