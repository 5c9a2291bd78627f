// C ABI of the engine (include/emc.h): context, host tables, device residency,
// initial population and the step orchestration on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/emc.h"
#include "band.h"
#include "device.h"
#include "fly_core.h"
#include "kernels.h"
#include "multigrid.h"
#include "noalloc.h"
#include "output.h"
#include "qc.h"
#include "tetindex.h"

using namespace emc;

struct DevBand;

namespace {

struct HostParticles {
  std::vector<double> x, y, z, kx, ky, kz, energy, charge;
  std::vector<int32_t> i, j, k, itet, isym, ptype;
  std::vector<uint64_t> key;
  size_t size() const { return x.size(); }
  void resize(size_t n) {
    for (auto* v : {&x, &y, &z, &kx, &ky, &kz, &energy, &charge}) v->resize(n);
    for (auto* v : {&i, &j, &k, &itet, &isym, &ptype}) v->resize(n);
    key.resize(n);
  }
};

struct DevBuf {
  std::vector<void*> owned;
  hipError_t last = hipSuccess;
  template <class T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) { last = e; return nullptr; }
    owned.push_back(p);
    return (T*)p;
  }
  template <class T>
  T* dup(const T* h, size_t n) {
    T* d = alloc<T>(n);
    if (d && n) {
      hipError_t e = hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice);
      if (e != hipSuccess) last = e;
    }
    return d;
  }
  template <class T>
  T* dup(const std::vector<T>& v) { return dup(v.data(), v.size()); }
  // free one allocation made here (the device must no longer use it)
  void release(const void* p) {
    if (!p) return;
    auto it = std::find(owned.begin(), owned.end(), p);
    if (it == owned.end()) return;
    (void)hipFree(*it);
    owned.erase(it);
  }
  void free_all() {
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
  }
};

struct Gpu {
  int device = -1;
  hipStream_t stream = nullptr;
  DevBuf mem;
  std::shared_ptr<DevBand> band_dev;  // the band tables on this GPU (shared)
  BandDev B{};
  MeshDev M{};
  FieldsDev F{};
  TallyDev T{};
  FlyConst* cst = nullptr;            // device copy of {B, M, F} for the flight kernels
  FlyConst cst_host{};                // what cst holds
  PartDev P{};
  PartDev P2{};                       // second particle buffer: the tile sort writes into it
  long long n = 0, cap = 0;
  // tile sort / deposit
  int ntx = 0, nty = 0, ntz = 0, ntile = 0;
  unsigned *tcount = nullptr, *toff = nullptr, *tcursor = nullptr, *tstay = nullptr, *toff_old = nullptr;
  // the population is in the tile segments of the last sort (toff), with the
  // particles added since after them from slot sorted_n on
  bool seg_valid = false;
  long long sorted_n = 0;
  unsigned long long* blk = nullptr;
  unsigned* n_host = nullptr;         // pinned: live count after the last sort
  hipEvent_t n_ready = nullptr;
  bool n_pending = false;
  CloneRec* clones = nullptr;
  unsigned long long clone_cap = 0;
  long long* dead = nullptr;
  unsigned long long dead_cap = 0;
  unsigned long long* ctr = nullptr;  // [0] nclones [1] ndead [2] ticket [3] slow-list length
                                      // [4] [5] clone range of the next pass [6] clone passes with work
  unsigned long long* ctr_host = nullptr;    // pinned copy of ctr after the flight
  unsigned long long* tally_host = nullptr;  // pinned copy of contact + counters after the flight
  bool pcg_pending = false;           // the step's PCG result is still to be read
  const double* pcg_result = nullptr; // where it will be: the pinned copy of the last queued iteration
  // MultipleRefresh work arrays, allocated on first use
  MrArgs mr{};
  bool mr_ready = false;
  long long mr_cap = 0;               // capacity the slot-sized MultipleRefresh arrays hold
  uint32_t* slow = nullptr;           // [cap] flight slow list
  // fields / mesh work arrays
  double *ex = nullptr, *ey = nullptr, *ez = nullptr, *rho = nullptr, *rates = nullptr;
  unsigned long long *node_q = nullptr, *node_stat = nullptr, *cell_q = nullptr, *cell_n = nullptr;
  unsigned long long *heat_e = nullptr, *heat_h = nullptr, *rcur = nullptr, *lcur = nullptr;
  unsigned long long *contact = nullptr, *counters = nullptr;
  // statistics window (statistic()'s sums since the last reset)
  double* stat_pot = nullptr;
  unsigned long long *stat_enum = nullptr, *stat_heat_e = nullptr, *stat_heat_h = nullptr;
  bool heat_live = false;             // heat was tallied since the last reset
  double* pot_avg = nullptr;          // heat phase: the fixed potential of the field
  // quantum correction (decks with Flag_QuantumCorrection)
  QcArgs qa{};                        // device tables
  double *qc_pot = nullptr, *stat_qc_pot = nullptr, *qc_avg = nullptr;
  double *hx = nullptr, *hy = nullptr, *hz = nullptr;  // hole fields of the Poisson potential
  // Poisson
  double *diag = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr, *dirichlet = nullptr, *corr = nullptr;
  double *pvol = nullptr, *pdop = nullptr, *rhs = nullptr, *pot = nullptr;
  double *r = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *z = nullptr, *scal = nullptr;
  double *part_a = nullptr, *part_pq = nullptr, *part_rr = nullptr, *part_rz0 = nullptr, *part_rz1 = nullptr;
  int* done = nullptr;
  MgArgs mg{};                        // device view of the multigrid hierarchy
  // FP32 copies for the bulk of the mixed-precision solve, made on its first use:
  // the level-0 operator, the hierarchy's vectors, damped inverse diagonals and
  // coarsest inverse, and the inner CG's vectors (its b is the FP32 residual, its
  // x the correction)
  bool f32_ready = false;
  PcgArgsT<float> fa{};
  MgArgsT<float> fmg{};
  // the FP32 cycle as the preconditioner of the FP64 CG: its input r32 (the
  // residual in FP32), its output z32, and the hierarchy above with level 0
  // reading r32 and the FP64 CG's done flag
  float *r32 = nullptr, *z32 = nullptr;
  MgArgsT<float> pmg{};
  double* scal_host = nullptr;        // pinned [2][8]: r.r etc. after iterations (ping-pong)
  hipEvent_t scal_ready[2] = {nullptr, nullptr};
  uint8_t* kind = nullptr;
  const double* cvol = nullptr;
  const double* dx = nullptr;
  const double* dy = nullptr;
  const double* dz = nullptr;
  const double* da = nullptr;
  const int* eeff = nullptr;
  int pcg_blocks = 0;
  long long last_full_loop = 0;       // particles of the last flight that took the full event loop
  long long wave_slots = 1024;        // CUs x 4 SIMDs
  hipEvent_t ev[8] = {};
};

// Particle arrays of one buffer, allocated together.
static void alloc_part(DevBuf& m, PartDev& P, long long cap) {
  P.x = m.alloc<double>(cap); P.y = m.alloc<double>(cap); P.z = m.alloc<double>(cap);
  P.kx = m.alloc<double>(cap); P.ky = m.alloc<double>(cap); P.kz = m.alloc<double>(cap);
  P.energy = m.alloc<double>(cap); P.charge = m.alloc<double>(cap);
  P.pk = m.alloc<uint32_t>(cap); P.itet = m.alloc<int32_t>(cap); P.isym = m.alloc<int32_t>(cap);
  P.key = m.alloc<uint64_t>(cap);
}

}  // namespace

// Per-cell contents of a restart file (read_par_info_for_restart).
struct RestartCells {
  std::vector<double> ec, hc;
  std::vector<int> en, hn;
  std::vector<uint8_t> have;
};

// The full-band tables and the records built from them: they depend only on
// the table files, the temperature and (quantum correction) the dispersive
// phonon files, so contexts of one process that agree on those share one copy,
// on the host and, per GPU, in device memory.
struct BandBundle {
  Band band;
  std::vector<TetRec> tetrec;
  std::vector<SurfRec> surfrec;
  std::vector<CornerRec> cornerrec;
  std::vector<Sphere> tsph, tlist_s, tlists_s, slist_s;
  TetIndex tix;         // empty when build_tet_index refused the tables
  std::string tix_why;  // why it refused
};

struct emc_ctx {
  std::string err;
  emc_opts opts{};
  Device dev;
  std::shared_ptr<const BandBundle> bb;
  std::vector<int> mr_desired_e, mr_desired_h;  // [ncell] MultipleRefresh target counts
  std::vector<int> surface_dir;
  HostParticles hp;
  std::unique_ptr<Gpu> gpu;
  MgHierarchy mg;                     // built on first use
  int last_pcg_iters = 0;
  double last_relres = 0;
  double timing[5] = {0, 0, 0, 0, 0};
  uint64_t next_step = 0;
  // statistics window: host side of the sums, and what output_stat keeps
  emc_fly_tallies win{};
  int64_t win_steps = 0;
  SavedStat saved;
  RestartCells restart;               // Flag_restart: the deck's restart file
  QcTables qc;                        // quantum-correction tables (Flag_QuantumCorrection)
  bool heat_phase = false;            // the reference's flag_heat
  uint64_t heat_next = 0;             // next heat-phase step
};

// The flight's output block: contact columns, counters, clone/merge counters.
static constexpr size_t FLYOUT_N = EMC_MAX_CONTACT * EMC_CT_NCOL + CNT_N + 8;

// Heat-phase step s draws from RNG step HEAT_STEP0 + s, apart from the transient's.
static constexpr uint64_t HEAT_STEP0 = 1ull << 40;
extern "C" {
static int run_steps(emc_ctx* c, int nsteps, emc_fly_tallies* acc, bool heat_phase);
}

static int fail(emc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHECK(ctx, expr)                                                              \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, EMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
  } while (0)

// ---------------------------------------------------------------- host band view
static void fill_band_dev_host(emc_ctx* c, BandDev& B) {
  const Band& b = c->bb->band;
  B.tr = c->bb->tetrec.data();
  B.sr = c->bb->surfrec.data();
  B.cr = c->bb->cornerrec.data();
  B.tsph = c->bb->tsph.data();
  B.troot = c->bb->tix.nodes.empty() ? nullptr : c->bb->tix.roots.data();
  B.tnode = c->bb->tix.nodes.empty() ? nullptr : c->bb->tix.nodes.data();
  B.tent = c->bb->tix.nodes.empty() ? nullptr : c->bb->tix.ent.data();
  B.tlist_s = c->bb->tlist_s.data(); B.tlists_s = c->bb->tlists_s.data(); B.slist_s = c->bb->slist_s.data();
  B.tet = b.tet.data();
  B.xkk = b.xkk.data(); B.ykk = b.ykk.data(); B.zkk = b.zkk.data(); B.eek = b.eek.data();
  B.vgt = b.vgt.data(); B.faclow = b.faclow.data(); B.facmed = b.facmedium.data(); B.fachigh = b.fachigh.data();
  B.matsym = &b.matsym[0][0]; B.indmat = b.indmat; B.newisym = &b.newisym[0][0];
  B.xbz = &b.xbz[0][0]; B.ybz = &b.ybz[0][0]; B.zbz = &b.zbz[0][0];
  B.ntlist = b.ntlist.data(); B.ptlist = b.ptlist.data(); B.tlist = b.tlist.data();
  B.nslist = b.nslist.data(); B.pslist = b.pslist.data(); B.slist = b.slist.data();
  B.ntlists = b.ntlists.data(); B.ptlists = b.ptlists.data(); B.tlists = b.tlists.data();
  B.ibt = b.ibt.data();
  B.energy = b.energy.data(); B.dostetmax = b.dostetmax.data(); B.sumscatt = b.sumscatt.data();
  B.dos = b.dos.data(); B.sumdos = b.sumdos.data();
  for (int q = 0; q < 3; ++q) B.dosmax[q] = b.DOSMAX[q];
  B.scatte = b.scatte.data(); B.dose = b.dose.data(); B.scattiie = b.scattiie.data();
  B.scatth = b.scatth.data(); B.dosh = b.dosh.data(); B.scattiih = b.scattiih.data();
  B.nt = b.nt; B.scpre = b.scpre; B.scprh = b.scprh;
  B.nbandE = b.nband[PELEC]; B.nbandH = b.nband[PHOLE];
  B.bandofE = b.bandof[PELEC]; B.bandofH = b.bandof[PHOLE];
  B.dtable = b.dtable; B.emax = b.emax; B.dlist = b.dlist; B.dlists = b.dlists;
  B.a0pi = b.a0pi; B.sieg = b.sieg; B.mell = b.mell; B.melt = b.melt; B.meld = b.meld;
  for (int q = 0; q < 3; ++q) B.difpr[q] = b.difpr[q];
  B.temptag = b.temptag; B.templag = b.templag; B.templog = b.templog; B.temptaf = b.temptaf;
  B.templaf = b.templaf; B.temptof = b.temptof; B.temphop = b.temphop;
  for (int t = 0; t < PTAB_N; ++t) B.ptab[t] = b.ptab_t[t].empty() ? nullptr : b.ptab_t[t].data();
}

static void fill_mesh_dev_scalars(emc_ctx* c, MeshDev& M) {
  const Device& d = c->dev;
  M.cnx = d.cnx; M.cny = d.cny; M.cnz = d.cnz; M.pnx = d.pnx; M.pny = d.pny; M.pnz = d.pnz;
  M.ncontact = (int)d.contacts.size();
  M.flag_surface = d.cfg.flag_surface; M.flag_simp = d.cfg.flag_simp;
  M.dt = d.cfg.dt; M.eV0 = d.phys.eV0; M.conc0 = d.phys.conc0; M.scrt0 = d.phys.scrt0;
  M.eps_si = d.phys.eps[SILICON]; M.frickel = d.phys.frickel;
  for (int q = 0; q < 3; ++q) M.charge_sign[q] = d.phys.charge_sign[q];
}

static void build_tetrec(BandBundle& bb) {
  const Band& b = bb.band;
  bb.tetrec.assign(b.nt, TetRec{});
  for (int it = 0; it < b.nt; ++it) {
    TetRec& r = bb.tetrec[it];
    for (int f = 0; f < 4; ++f)
      for (int q = 0; q < 4; ++q) r.n[f][q] = b.dn(q, f, it);
    for (int q = 0; q < 4; ++q) r.vt[q] = b.vt[4 * (size_t)it + q];
    r.gam = b.gamtet[it];
    r.e3 = b.eek[b.tet_corner(3, it)];
    for (int f = 0; f < 4; ++f) r.neib[f] = b.neib[(size_t)f * b.nt + it];
    r.ibt = b.ibt[it];
  }
  bb.surfrec.assign(b.nt, SurfRec{});
  bb.cornerrec.assign(b.nt, CornerRec{});
  for (int it = 0; it < b.nt; ++it) {
    SurfRec& s = bb.surfrec[it];
    CornerRec& q = bb.cornerrec[it];
    for (int k = 0; k < 4; ++k) {
      const int ck = b.tet_corner(k, it);
      s.e[k] = q.e[k] = b.eek[ck];
      q.x[k] = b.xkk[ck]; q.y[k] = b.ykk[ck]; q.z[k] = b.zkk[ck];
    }
    s.faclow = b.faclow[it]; s.facmed = b.facmedium[it]; s.fachigh = b.fachigh[it]; s.vgt = b.vgt[it];
  }
  // Enclosing spheres of the tetrahedra grown by the containment tolerance: corner
  // i of the grown tetrahedron solves n_f . v = d_f - MY_ZERO on the three faces
  // f != i (face f is opposite corner f).  Centre = centroid, radius = farthest
  // grown corner, widened by a relative 1e-9 for rounding.  A degenerate
  // tetrahedron gets an infinite radius (always tested exactly).
  // The box of the grown corners, widened the same way, bounds the tetrahedron for
  // the point-location index.
  bb.tsph.assign(b.nt, Sphere{});
  std::vector<TetBox> tbox(b.nt);
  for (int it = 0; it < b.nt; ++it) {
    double blo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, bhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    const CornerRec& q = bb.cornerrec[it];
    const double cx = 0.25 * (q.x[0] + q.x[1] + q.x[2] + q.x[3]);
    const double cy = 0.25 * (q.y[0] + q.y[1] + q.y[2] + q.y[3]);
    const double cz = 0.25 * (q.z[0] + q.z[1] + q.z[2] + q.z[3]);
    double r2max = 0.0;
    bool ok = true;
    for (int i = 0; i < 4 && ok; ++i) {
      double m[3][4];
      int row = 0;
      for (int f = 0; f < 4; ++f) {
        if (f == i) continue;
        for (int k = 0; k < 3; ++k) m[row][k] = b.dn(k, f, it);
        m[row][3] = b.dn(3, f, it) - MY_ZERO;
        ++row;
      }
      const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
      if (!(std::fabs(det) > 0.0)) { ok = false; break; }
      auto col = [&](int k, double& v) {
        double t[3][3];
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) t[r][cc] = cc == k ? m[r][3] : m[r][cc];
        v = (t[0][0] * (t[1][1] * t[2][2] - t[1][2] * t[2][1]) - t[0][1] * (t[1][0] * t[2][2] - t[1][2] * t[2][0]) +
             t[0][2] * (t[1][0] * t[2][1] - t[1][1] * t[2][0])) / det;
      };
      double vx, vy, vz;
      col(0, vx); col(1, vy); col(2, vz);
      const double r2 = (vx - cx) * (vx - cx) + (vy - cy) * (vy - cy) + (vz - cz) * (vz - cz);
      if (!std::isfinite(r2)) { ok = false; break; }
      if (r2 > r2max) r2max = r2;
      const double v[3] = {vx, vy, vz};
      for (int a = 0; a < 3; ++a) {
        blo[a] = std::min(blo[a], v[a]);
        bhi[a] = std::max(bhi[a], v[a]);
      }
    }
    const double r = ok ? std::sqrt(r2max) * (1.0 + 1e-9) + 1e-12 : HUGE_VAL;
    bb.tsph[it] = Sphere{cx, cy, cz, ok ? r * r : HUGE_VAL};
    for (int a = 0; a < 3; ++a) {
      const double w = 1e-9 * (std::fabs(blo[a]) + std::fabs(bhi[a]) + (bhi[a] - blo[a])) + 1e-12;
      tbox[it].lo[a] = ok ? blo[a] - w : -HUGE_VAL;
      tbox[it].hi[a] = ok ? bhi[a] + w : HUGE_VAL;
    }
  }
  auto along = [&](const std::vector<int>& list, std::vector<Sphere>& out) {
    out.resize(list.size());
    for (size_t q = 0; q < list.size(); ++q) out[q] = bb.tsph[list[q]];
  };
  along(b.tlist, bb.tlist_s);
  along(b.tlists, bb.tlists_s);
  along(b.slist, bb.slist_s);
  bb.tix_why = build_tet_index(b, tbox, bb.tix);
  if (!bb.tix_why.empty()) bb.tix = TetIndex{};
}

// Multigrid hierarchy of the Poisson preconditioner, built once per context.
// The band bundle for these tables and this temperature: built once per process
// and shared by every context that asks for the same one while any holds it.
static std::string band_bundle(const Device& d, const std::string& tdir, const std::string files[PTAB_N],
                               std::shared_ptr<const BandBundle>& out) {
  static std::mutex mu;
  static std::map<std::string, std::weak_ptr<const BandBundle>> cache;
  char tk[64];
  std::snprintf(tk, sizeof(tk), "%.17g", d.phys.T0);
  std::string key = tdir + "|" + tk;
  if (d.cfg.flag_qc)
    for (int t = 0; t < PTAB_N; ++t) key += "|" + files[t];
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end())
    if ((out = it->second.lock())) return "";
  auto bb = std::make_shared<BandBundle>();
  std::string err = band_build(bb->band, d.phys, tdir);
  if (!err.empty()) return err;
  if (d.cfg.flag_qc) {
    err = load_phonon_tables(bb->band, tdir, files);
    if (!err.empty()) return err;
  }
  build_tetrec(*bb);
  cache[key] = bb;
  out = bb;
  return "";
}

static std::string ensure_mg(emc_ctx* c) {
  if (!c->mg.lv.empty()) return "";
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // The coarsest level is solved by a dense inverse on the GPU (one wave per row):
  // about n0/300 nodes, between 600 and 2500, trades the launches of small levels
  // against the size of that product (1617 nodes and four coarse levels at the
  // 639 k-node FinFET mesh).
  const long long n0 = (long long)c->dev.nnode;
  const int min_nodes = (int)std::max(600LL, std::min(2500LL, n0 / 300));
  return mg_build(c->dev, c->mg, min_nodes, nth);
}

// ---------------------------------------------------------------- GPU setup
// The device copy of a band bundle on one GPU, shared by the contexts on that
// GPU that use the bundle; freed with the last of them.
struct DevBand {
  int device = -1;
  BandDev B{};
  DevBuf mem;
  std::shared_ptr<const BandBundle> bb;  // keeps the host tables the key points at
  ~DevBand() {
    int prev = -1;
    if (hipGetDevice(&prev) == hipSuccess && prev != device) (void)hipSetDevice(device);
    mem.free_all();
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
  }
};

static std::string device_band(emc_ctx* c, std::shared_ptr<DevBand>& out) {
  static std::mutex mu;
  static std::map<std::pair<int, const BandBundle*>, std::weak_ptr<DevBand>> cache;
  const int dev = c->opts.device;
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, c->bb.get());
  auto it = cache.find(key);
  if (it != cache.end())
    if ((out = it->second.lock())) return "";
  auto db = std::make_shared<DevBand>();
  db->device = dev;
  db->bb = c->bb;
  const BandBundle& bb = *c->bb;
  const Band& b = bb.band;
  DevBuf& m = db->mem;
  BandDev& B = db->B;
  fill_band_dev_host(c, B);  // scalars; pointers replaced below
  B.tr = m.dup(bb.tetrec);
  B.sr = m.dup(bb.surfrec);
  B.cr = m.dup(bb.cornerrec);
  B.tsph = m.dup(bb.tsph);
  if (!bb.tix.nodes.empty()) {
    B.troot = m.dup(bb.tix.roots);
    B.tnode = m.dup(bb.tix.nodes);
    B.tent = m.dup(bb.tix.ent);
  }
  B.tlist_s = m.dup(bb.tlist_s); B.tlists_s = m.dup(bb.tlists_s); B.slist_s = m.dup(bb.slist_s);
  B.tet = m.dup(b.tet);
  B.xkk = m.dup(b.xkk); B.ykk = m.dup(b.ykk); B.zkk = m.dup(b.zkk); B.eek = m.dup(b.eek);
  B.vgt = m.dup(b.vgt); B.faclow = m.dup(b.faclow); B.facmed = m.dup(b.facmedium); B.fachigh = m.dup(b.fachigh);
  B.matsym = m.dup(&b.matsym[0][0], 6 * 48); B.indmat = m.dup(b.indmat, 729);
  B.newisym = m.dup(&b.newisym[0][0], 5 * 48);
  B.xbz = m.dup(&b.xbz[0][0], 5 * 48); B.ybz = m.dup(&b.ybz[0][0], 5 * 48); B.zbz = m.dup(&b.zbz[0][0], 5 * 48);
  B.ntlist = m.dup(b.ntlist); B.ptlist = m.dup(b.ptlist); B.tlist = m.dup(b.tlist);
  B.nslist = m.dup(b.nslist); B.pslist = m.dup(b.pslist); B.slist = m.dup(b.slist);
  B.ntlists = m.dup(b.ntlists); B.ptlists = m.dup(b.ptlists); B.tlists = m.dup(b.tlists);
  B.ibt = m.dup(b.ibt);
  B.energy = m.dup(b.energy); B.dostetmax = m.dup(b.dostetmax); B.sumscatt = m.dup(b.sumscatt);
  B.dos = m.dup(b.dos); B.sumdos = m.dup(b.sumdos);
  for (int t = 0; t < PTAB_N; ++t) B.ptab[t] = b.ptab_t[t].empty() ? nullptr : m.dup(b.ptab_t[t]);
  B.scatte = m.dup(b.scatte); B.dose = m.dup(b.dose); B.scattiie = m.dup(b.scattiie);
  B.scatth = m.dup(b.scatth); B.dosh = m.dup(b.dosh); B.scattiih = m.dup(b.scattiih);
  if (m.last != hipSuccess) return std::string("band tables: ") + hipGetErrorString(m.last);
  cache[key] = db;
  out = db;
  return "";
}

static int gpu_setup(emc_ctx* c) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->opts.device)
    return fail(c, EMC_ERR_NOGPU, "no HIP device available for the requested ordinal");
  c->gpu.reset(new Gpu());
  Gpu& g = *c->gpu;
  g.device = c->opts.device;
  HIPCHECK(c, hipSetDevice(g.device));
  HIPCHECK(c, hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g.device) == hipSuccess && ncu > 0)
      g.wave_slots = 4LL * ncu;
  }
  for (auto& e : g.ev) HIPCHECK(c, hipEventCreate(&e));
  const Device& d = c->dev;
  DevBuf& m = g.mem;
  {
    const std::string e = device_band(c, g.band_dev);
    if (!e.empty()) return fail(c, EMC_ERR_HIP, e);
    g.B = g.band_dev->B;
  }

  MeshDev& M = g.M;
  fill_mesh_dev_scalars(c, M);
  M.lx = m.dup(d.lx); M.ly = m.dup(d.ly); M.lz = m.dup(d.lz);
  M.dx = m.dup(d.dx); M.dy = m.dup(d.dy); M.dz = m.dup(d.dz);
  M.motion = m.dup(d.motion); M.attached = m.dup(d.attached); M.insurf = m.dup(d.insurf);
  M.nearest_surf = m.dup(d.nearest_surf); M.da = m.dup(d.da);
  M.surface_dir = m.dup(c->surface_dir);
  g.dx = M.dx; g.dy = M.dy; g.dz = M.dz; g.da = M.da;
  g.cvol = m.dup(d.cvol);
  g.eeff = m.dup(d.eeff_dir);

  const size_t NC = d.ncell, NN = d.nnode;
  g.ex = m.alloc<double>(NC); g.ey = m.alloc<double>(NC); g.ez = m.alloc<double>(NC);
  g.rho = m.alloc<double>(NC); g.rates = m.alloc<double>(4 * NC);
  g.F.ex = g.ex; g.F.ey = g.ey; g.F.ez = g.ez;
  g.F.hx = g.ex; g.F.hy = g.ey; g.F.hz = g.ez;  // QC off: holes see the same field (src/mcmodel.cpp:2247-2262)
  g.F.rho = g.rho;
  g.F.e_rough = g.rates; g.F.e_sph = g.rates + NC; g.F.h_rough = g.rates + 2 * NC; g.F.h_sph = g.rates + 3 * NC;
  g.node_q = m.alloc<unsigned long long>(NN);
  g.node_stat = m.alloc<unsigned long long>(10 * NN);
  g.cell_q = m.alloc<unsigned long long>(3 * NC);
  g.cell_n = m.alloc<unsigned long long>(NC);
  g.heat_e = m.alloc<unsigned long long>(NN);
  g.heat_h = m.alloc<unsigned long long>(NN);
  g.stat_pot = m.alloc<double>(NN);
  g.stat_enum = m.alloc<unsigned long long>(NC);
  g.stat_heat_e = m.alloc<unsigned long long>(NN);
  g.stat_heat_h = m.alloc<unsigned long long>(NN);
  g.rcur = m.alloc<unsigned long long>(d.pny);
  g.lcur = m.alloc<unsigned long long>(d.pny);
  // contact columns, counters and the clone/merge counters in one block, so the
  // host reads all of them after the flight with one copy
  g.contact = m.alloc<unsigned long long>(FLYOUT_N);
  g.counters = g.contact + EMC_MAX_CONTACT * EMC_CT_NCOL;
  g.ctr = g.counters + CNT_N;
  g.cst = m.alloc<FlyConst>(1);
  g.ntx = (d.cnx + TILE_I - 1) / TILE_I;
  g.nty = (d.cny + TILE_J - 1) / TILE_J;
  g.ntz = (d.cnz + TILE_K - 1) / TILE_K;
  g.ntile = g.ntx * g.nty * g.ntz;
  g.tcount = m.alloc<unsigned>(g.ntile);
  g.tcursor = m.alloc<unsigned>(g.ntile);
  g.toff = m.alloc<unsigned>(g.ntile + 1);
  g.tstay = m.alloc<unsigned>(g.ntile);
  g.toff_old = m.alloc<unsigned>(g.ntile + 1);
  g.blk = m.alloc<unsigned long long>((size_t)g.ntile * DEP_NQ * TILE_NODES);
  g.T.contact = g.contact; g.T.counters = g.counters; g.T.rcur = g.rcur; g.T.lcur = g.lcur;
  g.T.heat_e = g.heat_e; g.T.heat_h = g.heat_h;

  g.diag = m.dup(d.diag); g.cx = m.dup(d.cx); g.cy = m.dup(d.cy); g.cz = m.dup(d.cz);
  g.dirichlet = m.dup(d.dirichlet_val); g.corr = m.dup(d.rhs_corr); g.kind = m.dup(d.node_kind);
  g.pvol = m.dup(d.pvol); g.pdop = m.dup(d.pdop);
  g.rhs = m.alloc<double>(NN); g.pot = m.alloc<double>(NN);
  g.r = m.alloc<double>(NN); g.p = m.alloc<double>(NN); g.p2 = m.alloc<double>(NN);
  g.q = m.alloc<double>(NN); g.z = m.alloc<double>(NN);
  g.pcg_blocks = emc_pcg_blocks(NN);
  for (double** pp : {&g.part_a, &g.part_pq, &g.part_rr, &g.part_rz0, &g.part_rz1}) *pp = m.alloc<double>(3 * g.pcg_blocks);
  g.scal = m.alloc<double>(8);
  g.done = m.alloc<int>(1);
  {
    const std::string e = ensure_mg(c);
    if (!e.empty()) return fail(c, EMC_ERR_DECK, e);
    const MgHierarchy& h = c->mg;
    if ((int)h.lv.size() > MG_MAXL) return fail(c, EMC_ERR_DECK, "mesh needs more multigrid levels than supported");
    MgArgs& A = g.mg;
    A.nl = (int)h.lv.size();
    A.nc = (int)h.lv.back().n;
    A.cinv = m.dup(h.cinv);
    A.diag = g.diag; A.cx = g.cx; A.cy = g.cy; A.cz = g.cz;
    A.done = g.done;
    for (int l = 0; l < A.nl; ++l) {
      const MgLevel& L = h.lv[l];
      MgLevelDev& D = A.L[l];
      D.nx = L.nx; D.ny = L.ny; D.nz = L.nz; D.n = L.n;
      D.c14 = nullptr;
      if (l > 0) {
        // upper half of the symmetric stencil (offsets 13..26 of st27) in FP32
        std::vector<float> c14(14 * L.n);
        for (int o = 13; o < 27; ++o)
          for (size_t q = 0; q < L.n; ++q) c14[(size_t)(o - 13) * L.n + q] = (float)L.a27[(size_t)o * L.n + q];
        D.c14 = m.dup(c14);
      }
      D.wdinv = m.dup(L.wdinv);
      D.x = m.alloc<double>(L.n); D.t = m.alloc<double>(L.n); D.r = m.alloc<double>(L.n);
      D.b = l > 0 ? m.alloc<double>(L.n) : g.r;
    }
  }
  if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(m.last));
  HIPCHECK(c, hipMemset(g.pot, 0, NN * sizeof(double)));
  HIPCHECK(c, hipMemset(g.tcount, 0, g.ntile * sizeof(unsigned)));
  HIPCHECK(c, hipMemset(g.toff, 0, (g.ntile + 1) * sizeof(unsigned)));
  HIPCHECK(c, hipHostMalloc((void**)&g.n_host, sizeof(unsigned), hipHostMallocDefault));
  HIPCHECK(c, hipHostMalloc((void**)&g.scal_host, 32 * sizeof(double), hipHostMallocDefault));
  HIPCHECK(c, hipHostMalloc((void**)&g.tally_host, FLYOUT_N * sizeof(unsigned long long), hipHostMallocDefault));
  g.ctr_host = g.tally_host + EMC_MAX_CONTACT * EMC_CT_NCOL + CNT_N;
  for (auto& e : g.scal_ready) HIPCHECK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHECK(c, hipEventCreateWithFlags(&g.n_ready, hipEventDisableTiming));
  HIPCHECK(c, hipMemset(g.ex, 0, NC * sizeof(double)));
  HIPCHECK(c, hipMemset(g.ey, 0, NC * sizeof(double)));
  HIPCHECK(c, hipMemset(g.ez, 0, NC * sizeof(double)));
  HIPCHECK(c, hipMemset(g.rates, 0, 4 * NC * sizeof(double)));
  std::vector<double> rho0(NC);
  for (size_t q = 0; q < NC; ++q) rho0[q] = 0.5 * d.da[q];
  HIPCHECK(c, hipMemcpy(g.rho, rho0.data(), NC * sizeof(double), hipMemcpyHostToDevice));
  for (auto* pz : {g.node_stat}) HIPCHECK(c, hipMemset(pz, 0, 10 * NN * 8));
  for (auto* pz : {g.heat_e, g.heat_h, g.stat_heat_e, g.stat_heat_h}) HIPCHECK(c, hipMemset(pz, 0, NN * 8));
  HIPCHECK(c, hipMemset(g.stat_pot, 0, NN * 8));
  HIPCHECK(c, hipMemset(g.stat_enum, 0, NC * 8));
  for (auto* pz : {g.rcur, g.lcur}) HIPCHECK(c, hipMemset(pz, 0, d.pny * 8));
  if (d.cfg.flag_qc) {
    const std::string e = build_qc_tables(d, c->qc);
    if (!e.empty()) return fail(c, EMC_ERR_DECK, e);
    const QcTables& q = c->qc;
    QcArgs& a = g.qa;
    a.nq = (long long)q.qnodes.size();
    a.pnx = d.pnx; a.pny = d.pny; a.pnz = d.pnz;
    a.qnodes = m.dup(q.qnodes);
    a.flag = m.dup(q.flag);
    for (int ax = 0; ax < 3; ++ax) {
      a.wlen[ax] = q.wlen[ax];
      a.lo[ax] = m.dup(q.lo[ax]);
      a.hi[ax] = m.dup(q.hi[ax]);
      a.w[ax] = m.dup(q.w[ax]);
    }
    a.eb = q.eb;
    g.qc_pot = m.alloc<double>(NN);
    g.stat_qc_pot = m.alloc<double>(NN);
    g.hx = m.alloc<double>(NC); g.hy = m.alloc<double>(NC); g.hz = m.alloc<double>(NC);
    if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(m.last));
    for (double* pz : {g.qc_pot, g.stat_qc_pot}) HIPCHECK(c, hipMemset(pz, 0, NN * 8));
    for (double* pz : {g.hx, g.hy, g.hz}) HIPCHECK(c, hipMemset(pz, 0, NC * 8));
    g.F.hx = g.hx; g.F.hy = g.hy; g.F.hz = g.hz;
  }
  HIPCHECK(c, hipDeviceSynchronize());
  return EMC_OK;
}

static void release_part(DevBuf& m, PartDev& P) {
  for (const void* p : {(const void*)P.x, (const void*)P.y, (const void*)P.z, (const void*)P.kx, (const void*)P.ky,
                        (const void*)P.kz, (const void*)P.energy, (const void*)P.charge, (const void*)P.pk,
                        (const void*)P.itet, (const void*)P.isym, (const void*)P.key})
    m.release(p);
  P = PartDev{};
}

static int gpu_alloc_particles(emc_ctx* c, long long cap) {
  Gpu& g = *c->gpu;
  if (g.cap >= cap) return EMC_OK;
  // particle slots are listed as 32-bit indices by the flight kernels
  if (cap > 0x7fffffffLL) return fail(c, EMC_ERR_CAPACITY, "particle capacity above 2^31-1");
  DevBuf& m = g.mem;
  if (g.cap > 0) {
    // a larger population replaces the current one (every caller writes all of
    // it): the old buffers go once no queued kernel reads them
    HIPCHECK(c, hipStreamSynchronize(g.stream));
    release_part(m, g.P);
    release_part(m, g.P2);
    m.release(g.clones); m.release(g.dead); m.release(g.slow);
    g.clones = nullptr; g.dead = nullptr; g.slow = nullptr;
    g.n = 0;
    g.n_pending = false;
  }
  g.seg_valid = false;
  g.cap = cap;
  alloc_part(m, g.P, cap);
  alloc_part(m, g.P2, cap);
  g.clone_cap = (unsigned long long)std::max<long long>(4096, cap / 16);
  g.clones = m.alloc<CloneRec>(g.clone_cap);
  g.dead_cap = g.clone_cap;
  g.dead = m.alloc<long long>(g.dead_cap);
  g.slow = m.alloc<uint32_t>(cap);
  if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("particle allocation: ") + hipGetErrorString(m.last));
  return EMC_OK;
}

// Keeps room for a step's clones: once the population comes within one clone
// buffer of its slots, the particles move into buffers of 1.5x the population
// (the reference's cell lists have no limit).  Between steps only: the flight
// appends a step's surviving clones up to the capacity and counts the rest as
// drops (reason 8).  The stream must have passed every kernel that reads the
// particles, and g.n must be settled.
static int grow_if_full(emc_ctx* c) {
  Gpu& g = *c->gpu;
  if (g.n <= g.cap - (long long)g.clone_cap) return EMC_OK;
  const long long cap = (long long)std::ceil(1.5 * (double)g.n) + 4096;
  if (cap > 0x7fffffffLL) return fail(c, EMC_ERR_CAPACITY, "particle capacity above 2^31-1");
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  DevBuf& m = g.mem;
  PartDev out{}, spare{};
  alloc_part(m, out, cap);
  alloc_part(m, spare, cap);
  const unsigned long long clone_cap = (unsigned long long)std::max<long long>(4096, cap / 16);
  CloneRec* clones = m.alloc<CloneRec>(clone_cap);
  long long* dead = m.alloc<long long>(clone_cap);
  uint32_t* slow = m.alloc<uint32_t>(cap);
  if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("particle allocation: ") + hipGetErrorString(m.last));
  const size_t n = (size_t)g.n;
  const std::pair<const void*, void*> f8[] = {{g.P.x, out.x}, {g.P.y, out.y}, {g.P.z, out.z}, {g.P.kx, out.kx},
                                              {g.P.ky, out.ky}, {g.P.kz, out.kz}, {g.P.energy, out.energy},
                                              {g.P.charge, out.charge}, {g.P.key, out.key}};
  for (const auto& pr : f8) HIPCHECK(c, hipMemcpyAsync(pr.second, pr.first, n * 8, hipMemcpyDeviceToDevice, g.stream));
  const std::pair<const void*, void*> f4[] = {{g.P.pk, out.pk}, {g.P.itet, out.itet}, {g.P.isym, out.isym}};
  for (const auto& pr : f4) HIPCHECK(c, hipMemcpyAsync(pr.second, pr.first, n * 4, hipMemcpyDeviceToDevice, g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  release_part(m, g.P);
  release_part(m, g.P2);
  m.release(g.clones); m.release(g.dead); m.release(g.slow);
  g.P = out;
  g.P2 = spare;
  g.cap = cap;
  g.clones = clones; g.clone_cap = clone_cap;
  g.dead = dead; g.dead_cap = clone_cap;
  g.slow = slow;
  return EMC_OK;
}

// MultipleRefresh target counts per cell (src/particle_init.cpp:44-70): the
// deck's per-cell numbers, raised in silicon cells to the initial particle
// count of the cell and to the default_par_number floor.
static void mr_targets(emc_ctx* c) {
  const Device& d = c->dev;
  c->mr_desired_e = d.desired_e;
  c->mr_desired_h = d.desired_h;
  if (d.sum_charge[0] == 0.0 || d.sum_charge[1] == 0.0) return;
  for (size_t cc = 0; cc < (size_t)d.ncell; ++cc) {
    if (d.material[cc] != SILICON) continue;
    const int ne = (int)(1e-10 + d.init_e_charge[cc] / d.sum_charge[0] * d.cfg.electron_number) + 1;
    const int nh = (int)(1e-10 + d.init_h_charge[cc] / d.sum_charge[1] * d.cfg.hole_number) + 1;
    int& de = c->mr_desired_e[cc];
    int& dh = c->mr_desired_h[cc];
    if (ne > de) de = ne;
    if (nh > dh) dh = nh;
    if (de < d.default_e) de = d.default_e;
    if (dh < d.default_h) dh = d.default_h;
  }
}

// ---------------------------------------------------------------- initial population
// init_particle_data / compute_par_num / select_kstate (src/particle_init.cpp:17-219,
// src/mcmodel.cpp:3006-3057), with the per-particle Philox stream of the build
// (draw stream step = 2^64-1 reserved for initialisation).  Particle s of the
// population is number s of the reference's creation order: type, then cells
// with i outer, k, j inner, then the cell's count; its key is init_key(seed, s).
struct SpecGroup { int i, j, k, type, count; double q; };

// read_par_info_for_restart (src/mcmodel.cpp:853-889): rows "i j k |Qe| Ne |Qh| Nh"
// up to a row with i < 0; rows of cells outside the mesh are skipped.
static std::string read_restart(const Device& d, const std::string& path, RestartCells& r) {
  std::ifstream f(path);
  if (!f) return "cannot open restart file " + path;
  r.ec.assign(d.ncell, 0.0); r.hc.assign(d.ncell, 0.0);
  r.en.assign(d.ncell, 0); r.hn.assign(d.ncell, 0);
  r.have.assign(d.ncell, 0);
  for (;;) {
    long long i, j, k;
    if (!(f >> i >> j >> k)) return "restart file " + path + ": no -1 -1 -1 terminator";
    if (i < 0) break;
    double c1, c2;
    long long n1, n2;
    if (!(f >> c1 >> n1 >> c2 >> n2)) return "restart file " + path + ": malformed row";
    if (n1 < 0 || n2 < 0 || n1 > INT32_MAX || n2 > INT32_MAX) return "restart file " + path + ": bad particle count";
    if (i >= d.cnx || j < 0 || j >= d.cny || k < 0 || k >= d.cnz) continue;
    const size_t cc = d.cell((int)i, (int)j, (int)k);
    r.ec[cc] = c1; r.en[cc] = (int)n1; r.hc[cc] = c2; r.hn[cc] = (int)n2; r.have[cc] = 1;
  }
  return "";
}

// compute_par_num's counts, or with a restart table (src/particle_init.cpp:128-170)
// the file's counts; cells attached to a contact keep their neutral charge, the
// others share the file's charge (electrons negative).
static std::vector<SpecGroup> spec_groups(const Device& d, const RestartCells* r) {
  std::vector<SpecGroup> gr;
  std::vector<int> ne(d.ncell, 0), nh(d.ncell, 0);
  for (int i = 0; i < d.cnx; ++i)
    for (int k = 0; k < d.cnz; ++k)
      for (int j = 0; j < d.cny; ++j) {
        size_t cc = d.cell(i, j, k);
        if (d.material[cc] != SILICON) continue;
        ne[cc] = (int)(1e-10 + d.init_e_charge[cc] / d.sum_charge[0] * d.cfg.electron_number) + 1;
        nh[cc] = (int)(1e-10 + d.init_h_charge[cc] / d.sum_charge[1] * d.cfg.hole_number) + 1;
        if (r && r->have[cc]) { ne[cc] = r->en[cc]; nh[cc] = r->hn[cc]; }
      }
  for (int t = 0; t < 2; ++t)
    for (int i = 0; i < d.cnx; ++i)
      for (int k = 0; k < d.cnz; ++k)
        for (int j = 0; j < d.cny; ++j) {
          size_t cc = d.cell(i, j, k);
          if (d.material[cc] != SILICON) continue;
          const int pn = t == 0 ? ne[cc] : nh[cc];
          if (pn == 0) continue;
          double q;
          if (!r || d.attached[cc] > 0) q = (t == 0 ? d.init_e_charge[cc] : d.init_h_charge[cc]) / pn;
          else q = t == 0 ? -r->ec[cc] / pn : r->hc[cc] / pn;
          gr.push_back(SpecGroup{i, j, k, t, pn, q});
        }
  return gr;
}

// The restart table of the deck (Flag_restart), read on first use.
static std::string restart_table(emc_ctx* c, const RestartCells** out) {
  *out = nullptr;
  const Config& cf = c->dev.cfg;
  if (!cf.flag_restart) return "";
  if (c->restart.have.empty()) {
    std::string path = cf.restart_file;
    if (!path.empty() && path[0] != '/' && !cf.dir.empty()) path = cf.dir + "/" + path;
    const std::string e = read_restart(c->dev, path, c->restart);
    if (!e.empty()) return e;
  }
  *out = &c->restart;
  return "";
}

static void sample_population(emc_ctx* c, const RestartCells* rs) {
  const Device& d = c->dev;
  const Band& b = c->bb->band;
  struct Spec { int i, j, k, type; double q; };
  std::vector<Spec> spec;
  for (const SpecGroup& g : spec_groups(d, rs))
    for (int a = 0; a < g.count; ++a) spec.push_back(Spec{g.i, g.j, g.k, g.type, g.q});
  HostParticles& hp = c->hp;
  const size_t n = spec.size();
  hp.resize(n);
  BandDev B{};
  fill_band_dev_host(c, B);
  MeshDev M{};
  fill_mesh_dev_scalars(c, M);
  FieldsDev F{};
  TallyDev T{};
  const double tmp_const = 1 - std::exp(-b.emax);
  const uint64_t seed = c->opts.seed;
  auto work = [&](size_t lo, size_t hi) {
    Flight f(B, M, F, T);
    for (size_t s = lo; s < hi; ++s) {
      const Spec& sp = spec[s];
      const uint64_t key = init_key(seed, s);
      f.rng.init(key, ~0ull);
      double lam = f.Random();
      hp.x[s] = lam * d.lx[sp.i] + (1 - lam) * d.lx[sp.i + 1];
      lam = f.Random();
      hp.y[s] = lam * d.ly[sp.j] + (1 - lam) * d.ly[sp.j + 1];
      lam = f.Random();
      hp.z[s] = lam * d.lz[sp.k] + (1 - lam) * d.lz[sp.k + 1];
      f.par_type = sp.type;
      f.sample_kstate(tmp_const);
      hp.kx[s] = f.kx; hp.ky[s] = f.ky; hp.kz[s] = f.kz; hp.energy[s] = f.energy;
      hp.charge[s] = sp.q;
      hp.i[s] = sp.i; hp.j[s] = sp.j; hp.k[s] = sp.k;
      hp.itet[s] = f.itet; hp.isym[s] = f.isym; hp.ptype[s] = sp.type; hp.key[s] = key;
    }
  };
  unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 20000) nth = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t) th.emplace_back(work, n * t / nth, n * (t + 1) / nth);
  for (auto& t : th) t.join();
}

static int settle_n(emc_ctx* c);

static int gpu_upload_host(emc_ctx* c) {
  Gpu& g = *c->gpu;
  const HostParticles& hp = c->hp;
  const long long n = (long long)hp.size();
  const long long want = (long long)std::ceil(n * std::max(1.0, c->opts.capacity_factor)) + 1024;
  int st = settle_n(c);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  st = gpu_alloc_particles(c, want);
  if (st) return st;
  if (n > g.cap) return fail(c, EMC_ERR_CAPACITY, "particle count exceeds the device capacity");
  std::vector<uint32_t> pk(n);
  for (long long s = 0; s < n; ++s)
    pk[s] = hp.i[s] < 0 ? PK_DEAD : pk_make(hp.ptype[s], hp.i[s], hp.j[s], hp.k[s]);
  auto up = [&](void* dptr, const void* h, size_t bytes) { return hipMemcpy(dptr, h, bytes, hipMemcpyHostToDevice); };
  HIPCHECK(c, up(g.P.x, hp.x.data(), n * 8)); HIPCHECK(c, up(g.P.y, hp.y.data(), n * 8));
  HIPCHECK(c, up(g.P.z, hp.z.data(), n * 8)); HIPCHECK(c, up(g.P.kx, hp.kx.data(), n * 8));
  HIPCHECK(c, up(g.P.ky, hp.ky.data(), n * 8)); HIPCHECK(c, up(g.P.kz, hp.kz.data(), n * 8));
  HIPCHECK(c, up(g.P.energy, hp.energy.data(), n * 8)); HIPCHECK(c, up(g.P.charge, hp.charge.data(), n * 8));
  HIPCHECK(c, up(g.P.pk, pk.data(), n * 4)); HIPCHECK(c, up(g.P.itet, hp.itet.data(), n * 4));
  HIPCHECK(c, up(g.P.isym, hp.isym.data(), n * 4)); HIPCHECK(c, up(g.P.key, hp.key.data(), n * 8));
  g.n = n;
  g.seg_valid = false;
  return EMC_OK;
}

static int sync_fly_const(emc_ctx* c);

// The same sampling on the GPU, one lane per particle (k_sample_population): the
// particles are written straight into the device population, identical to the
// host sampler's slot for slot.
static int gpu_sample_population(emc_ctx* c, const RestartCells* rs) {
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  const std::vector<SpecGroup> gr = spec_groups(c->dev, rs);
  std::vector<unsigned long long> start(gr.size() + 1, 0);
  std::vector<uint32_t> gpk(gr.size());
  std::vector<double> gq(gr.size());
  for (size_t q = 0; q < gr.size(); ++q) {
    start[q + 1] = start[q] + (unsigned long long)gr[q].count;
    gpk[q] = pk_make(gr[q].type, gr[q].i, gr[q].j, gr[q].k);
    gq[q] = gr[q].q;
  }
  const long long n = (long long)start.back();
  const long long want = (long long)std::ceil(n * std::max(1.0, c->opts.capacity_factor)) + 1024;
  st = gpu_alloc_particles(c, want);
  if (st) return st;
  if (n > g.cap) return fail(c, EMC_ERR_CAPACITY, "particle count exceeds the device capacity");
  st = sync_fly_const(c);
  if (st) return st;
  DevBuf tmp;
  SampleArgs a{};
  a.C = g.cst; a.P = g.P; a.n = n; a.ngroup = (long long)gr.size();
  a.start = tmp.dup(start); a.gpk = tmp.dup(gpk); a.gq = tmp.dup(gq);
  a.seed = c->opts.seed;
  a.tmp_const = 1 - std::exp(-c->bb->band.emax);
  if (tmp.last != hipSuccess) {
    tmp.free_all();
    return fail(c, EMC_ERR_HIP, std::string("sampling tables: ") + hipGetErrorString(tmp.last));
  }
  hipError_t e = emc_launch_sample(a, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  tmp.free_all();
  if (e != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("initial sampling: ") + hipGetErrorString(e));
  g.n = n;
  g.n_pending = false;
  g.seg_valid = false;
  c->hp.resize(0);
  return EMC_OK;
}

// ---------------------------------------------------------------- public API
static int ensure_f32(emc_ctx* c);

extern "C" {

void emc_default_opts(emc_opts* o) {
  if (!o) return;
  o->table_dir = nullptr;
  o->device = 0;
  o->seed = 1;
  o->capacity_factor = 1.25;
  o->poisson_rtol = 1e-10;
  o->poisson_maxit = 20000;
  o->poisson_fp32 = 2;
  o->poisson_inner_rtol = 1e-5;
}

emc_ctx* emc_create(const char* input_txt, const emc_opts* o, int* status) {
  emc_ctx* c = new emc_ctx();
  auto done = [&](int st) -> emc_ctx* {
    if (status) *status = st;
    return c;
  };
  if (!input_txt) return done(fail(c, EMC_ERR_ARG, "input path is NULL"));
  if (o) c->opts = *o;
  else emc_default_opts(&c->opts);
  if (c->opts.poisson_fp32 < 0 || c->opts.poisson_fp32 > 2)
    return done(fail(c, EMC_ERR_ARG, "poisson_fp32 must be 0 (FP64), 1 (FP32 bulk with FP64 refinement) or 2 (FP32 cycle)"));
  std::string err = device_parse(c->dev, input_txt);
  if (!err.empty()) return done(fail(c, err.find("cannot open") != std::string::npos ? EMC_ERR_IO : EMC_ERR_DECK, err));
  std::string tdir = c->opts.table_dir ? c->opts.table_dir : "";
  if (tdir.empty()) return done(fail(c, EMC_ERR_ARG, "table_dir is required"));
  {
    // EPSCAT takes optical and f-phonon final energies from the dispersive
    // tables in quantum-correction runs (src/ElectronPhononScatter.cpp:148-163)
    std::string files[PTAB_N] = {"kloem.txt", "kloab.txt", "ktoem.txt", "ktoab.txt", "klaem.txt", "klaab.txt"};
    const char* keys[PTAB_N] = {"kloem_file", "kloab_file", "ktoem_file", "ktoab_file", "klaem_file", "klaab_file"};
    for (int t = 0; t < PTAB_N; ++t) {
      auto it = c->dev.cfg.raw.find(keys[t]);
      if (it != c->dev.cfg.raw.end()) files[t] = it->second;
    }
    err = band_bundle(c->dev, tdir, files, c->bb);
    if (!err.empty()) return done(fail(c, EMC_ERR_IO, err));
  }
  err = device_finish(c->dev, c->bb->band);
  if (!err.empty()) return done(fail(c, EMC_ERR_DECK, err));
  c->surface_dir.assign(std::max(1, c->dev.nsurf), 0);
  for (int s = 0; s < c->dev.nsurf; ++s) c->surface_dir[s] = c->dev.surf_dir[s];
  mr_targets(c);
  if (c->opts.device >= 0) {
    int st = gpu_setup(c);
    if (st) return done(st);
    // the scratch array of emc_poisson_solve_noalloc: its device address, once
    // per device, with the device current
    st = emc_noalloc_setup(c->gpu->device);
    if (st) return done(fail(c, st, "emc_poisson_solve_noalloc: the scratch array's device address is not available"));
    // the FP32 cycle's operator and hierarchy, so that no solve allocates
    if (c->opts.poisson_fp32 == 2) {
      st = ensure_f32(c);
      if (st) return done(st);
    }
  }
  return done(EMC_OK);
}

void emc_destroy(emc_ctx* c) {
  if (!c) return;
  if (c->gpu) {
    (void)hipSetDevice(c->gpu->device);
    (void)hipDeviceSynchronize();
    for (auto e : c->gpu->ev)
      if (e) (void)hipEventDestroy(e);
    if (c->gpu->n_ready) (void)hipEventDestroy(c->gpu->n_ready);
    if (c->gpu->n_host) (void)hipHostFree(c->gpu->n_host);
    if (c->gpu->scal_host) (void)hipHostFree(c->gpu->scal_host);
    if (c->gpu->tally_host) (void)hipHostFree(c->gpu->tally_host);
    for (auto e : c->gpu->scal_ready)
      if (e) (void)hipEventDestroy(e);
    if (c->gpu->stream) (void)hipStreamDestroy(c->gpu->stream);
    c->gpu->mem.free_all();
  }
  delete c;
}

const char* emc_last_error(emc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int emc_band_view_get(emc_ctx* c, emc_band_view* v) {
  if (!c || !v) return EMC_ERR_ARG;
  const Band& b = c->bb->band;
  std::memset(v, 0, sizeof(*v));
  v->nk = b.nk; v->nt = b.nt; v->nbt = b.nbt;
  for (int q = 0; q < 3; ++q) { v->nband[q] = b.nband[q]; v->bandof[q] = b.bandof[q]; v->difpr[q] = b.difpr[q]; v->DOSMAX[q] = b.DOSMAX[q]; }
  v->tsphere = c->bb->tsph.empty() ? nullptr : &c->bb->tsph[0].cx;
  v->tindex_nodes = (int)c->bb->tix.nodes.size();
  v->tindex_entries = (int)c->bb->tix.ent.size();
  if (!c->bb->tix.nodes.empty()) {
    v->tindex_root = c->bb->tix.roots.data();
    v->tindex_node = &c->bb->tix.nodes[0].first;
    v->tindex_entry = c->bb->tix.ent.data();
  }
  v->scpre = b.scpre; v->scprh = b.scprh;
  v->dtable = b.dtable; v->emax = b.emax; v->dlist = b.dlist; v->dlists = b.dlists; v->a0pi = b.a0pi; v->sieg = b.sieg;
  v->mell = b.mell; v->melt = b.melt; v->meld = b.meld;
  v->temptag = b.temptag; v->templag = b.templag; v->templog = b.templog; v->temptaf = b.temptaf;
  v->templaf = b.templaf; v->temptof = b.temptof; v->temphop = b.temphop; v->Ni = b.Ni;
  v->xkk = b.xkk.data(); v->ykk = b.ykk.data(); v->zkk = b.zkk.data(); v->eek = b.eek.data();
  v->tet = b.tet.data(); v->neib = b.neib.data(); v->ibt = b.ibt.data(); v->dnt = b.dnt.data(); v->vt = b.vt.data();
  v->vgt = b.vgt.data(); v->gamtet = b.gamtet.data(); v->faclow = b.faclow.data(); v->facmedium = b.facmedium.data();
  v->fachigh = b.fachigh.data(); v->maxaovtet = b.maxaovtet.data();
  v->xkct = b.xkct.data(); v->ykct = b.ykct.data(); v->zkct = b.zkct.data();
  v->matsym = &b.matsym[0][0]; v->indmat = b.indmat; v->newisym = &b.newisym[0][0];
  v->xbz = &b.xbz[0][0]; v->ybz = &b.ybz[0][0]; v->zbz = &b.zbz[0][0];
  v->ntlist = b.ntlist.data(); v->ptlist = b.ptlist.data(); v->tlist = b.tlist.data();
  v->nslist = b.nslist.data(); v->pslist = b.pslist.data(); v->slist = b.slist.data();
  v->ntlists = b.ntlists.data(); v->ptlists = b.ptlists.data(); v->tlists = b.tlists.data();
  v->tlist_len = (int)b.tlist.size(); v->slist_len = (int)b.slist.size(); v->tlists_len = (int)b.tlists.size();
  v->energy = b.energy.data(); v->dos = b.dos.data(); v->sumdos = b.sumdos.data(); v->dostetmax = b.dostetmax.data();
  v->sumscatt = b.sumscatt.data(); v->scatte = b.scatte.data(); v->dose = b.dose.data(); v->scattiie = b.scattiie.data();
  v->scatth = b.scatth.data(); v->dosh = b.dosh.data(); v->scattiih = b.scattiih.data();
  for (int t = 0; t < PTAB_N; ++t) v->ptab[t] = b.ptab[t].empty() ? nullptr : b.ptab[t].data();
  return EMC_OK;
}

int emc_band_locate(emc_ctx* c, int64_t n, const double* k, const double* energy, const int* iband, int* itet,
                    int* isym) {
  if (!c) return EMC_ERR_ARG;
  if (n < 0 || (n > 0 && (!k || !energy || !iband || !itet || !isym))) return fail(c, EMC_ERR_ARG, "band_locate: bad arguments");
  const Band& b = c->bb->band;
  for (int64_t q = 0; q < n; ++q)
    if (iband[q] < 0 || iband[q] >= MNB || itet[q] < 0 || itet[q] >= b.nt)
      return fail(c, EMC_ERR_ARG, "band_locate: band or tetrahedron index out of range");
  BandDev B{};
  fill_band_dev_host(c, B);
  MeshDev M{};
  fill_mesh_dev_scalars(c, M);
  FieldsDev F{};
  TallyDev T{};
  Flight f(B, M, F, T);
  for (int64_t q = 0; q < n; ++q) {
    f.kx = k[3 * q]; f.ky = k[3 * q + 1]; f.kz = k[3 * q + 2];
    f.energy = energy[q];
    f.iband = iband[q];
    f.itet = itet[q];
    f.Flag_Catch = 0;
    f.Getisym();
    itet[q] = f.Flag_Catch ? -1 : f.itet;
    isym[q] = f.isym;
  }
  return EMC_OK;
}

int emc_band_final_state(emc_ctx* c, int64_t n, const uint64_t* key, uint64_t step, const double* energy,
                         const int* iband, const int* isym, double* k, int* itet, int* ndraws, int* selfsc,
                         int serial) {
  if (!c) return EMC_ERR_ARG;
  if (n < 0 || (n > 0 && (!key || !energy || !iband || !isym || !k || !itet || !ndraws || !selfsc)))
    return fail(c, EMC_ERR_ARG, "band_final_state: bad arguments");
  const Band& b = c->bb->band;
  for (int64_t q = 0; q < n; ++q)
    if (iband[q] < 0 || iband[q] >= MNB || isym[q] < 0 || isym[q] >= 48 || itet[q] < 0 || itet[q] >= b.nt)
      return fail(c, EMC_ERR_ARG, "band_final_state: band, symmetry or tetrahedron index out of range");
  BandDev B{};
  fill_band_dev_host(c, B);
  MeshDev M{};
  fill_mesh_dev_scalars(c, M);
  FieldsDev F{};
  TallyDev T{};
  Flight f(B, M, F, T);
  for (int64_t q = 0; q < n; ++q) {
    f.rng.init(key[q], step);
    f.energy = energy[q];
    f.iband = iband[q];
    f.isym = isym[q];
    f.itet = itet[q];
    f.kx = f.ky = f.kz = 0.0;
    f.Flag_SelfScatter = 0;
    if (serial < 0) f.get_state();
    else f.get_state_chunked(serial);
    k[3 * q] = f.kx; k[3 * q + 1] = f.ky; k[3 * q + 2] = f.kz;
    itet[q] = f.itet;
    ndraws[q] = (int)(f.rng.block * 2u - (f.rng.have ? 1u : 0u));
    selfsc[q] = f.Flag_SelfScatter;
  }
  return EMC_OK;
}

int emc_mesh_view_get(emc_ctx* c, emc_mesh_view* v) {
  if (!c || !v) return EMC_ERR_ARG;
  const Device& d = c->dev;
  const Phys& p = d.phys;
  std::memset(v, 0, sizeof(*v));
  v->cnx = d.cnx; v->cny = d.cny; v->cnz = d.cnz; v->pnx = d.pnx; v->pny = d.pny; v->pnz = d.pnz;
  v->ncell = (int64_t)d.ncell; v->nnode = (int64_t)d.nnode;
  v->ncontact = (int)d.contacts.size(); v->nsurf = d.nsurf;
  v->dt = d.cfg.dt;
  v->spr0 = p.spr0; v->time0 = p.time0; v->eV0 = p.eV0; v->pot0 = p.pot0; v->conc0 = p.conc0; v->scrt0 = p.scrt0;
  v->field0 = p.field0; v->curr0 = p.curr0; v->velo0 = p.velo0;
  v->eps_si = p.eps[SILICON]; v->eps_ox = p.eps[OXIDE]; v->frickel = p.frickel; v->Tn = p.Tn; v->T0 = p.T0;
  v->em0 = p.em0; v->hq0 = p.hq0; v->ec0 = p.ec0; v->sirho = p.sirho;
  v->surf_delta = p.SurfSc_delta; v->surf_ail = p.SurfSc_ail; v->surf_XIph = p.SurfSc_XIph;
  v->surf_theta = p.SurfSc_theta; v->surf_gama = p.SurfSc_gama; v->surf_Nbmod = p.SurfSc_Nbmod;
  v->surf_Npf = p.SurfSc_Npf; v->surf_Pft = p.SurfSc_Pft; v->surf_Pfn = p.SurfSc_Pfn; v->surf_Kpha = p.SurfSc_Kpha;
  for (int q = 0; q < 3; ++q) { v->mass_eff[q] = p.mass_eff[q]; v->charge_sign[q] = p.charge_sign[q]; }
  v->flag_surface = d.cfg.flag_surface; v->flag_rough = d.cfg.flag_rough; v->flag_sphonon = d.cfg.flag_sphonon;
  v->flag_simp = d.cfg.flag_simp; v->flag_heat = d.cfg.flag_heat;
  v->lx = d.lx.data(); v->ly = d.ly.data(); v->lz = d.lz.data(); v->dx = d.dx.data(); v->dy = d.dy.data(); v->dz = d.dz.data();
  v->material = d.material.data(); v->attached = d.attached.data(); v->nearest_surf = d.nearest_surf.data();
  v->eeff_dir = d.eeff_dir.data(); v->motion = d.motion.data(); v->insurf = d.insurf.data();
  v->cvol = d.cvol.data(); v->da = d.da.data(); v->surface_dir = c->surface_dir.data();
  v->pvol = d.pvol.data(); v->pdop = d.pdop.data(); v->vadd = d.vadd.data();
  v->node_kind = d.node_kind.data(); v->dirichlet = d.dirichlet_val.data();
  v->diag = d.diag.data(); v->cx = d.cx.data(); v->cy = d.cy.data(); v->cz = d.cz.data();
  v->rhs_corr = d.rhs_corr.data(); v->stencil = d.stencil.data();
  v->mr_desired_e = c->mr_desired_e.empty() ? nullptr : c->mr_desired_e.data();
  v->mr_desired_h = c->mr_desired_h.empty() ? nullptr : c->mr_desired_h.data();
  v->flag_qc = d.cfg.flag_qc; v->quantum_start_step = d.cfg.quantum_start_step;
  for (int a = 0; a < 3; ++a) { v->qc_ratio[a] = d.qc_ratio[a]; v->qc_theta[a] = d.qc_theta[a]; }
  v->qc_eb = d.qc_eb;
  v->qcregion = d.qcregion.data();
  return EMC_OK;
}

int emc_particles_init(emc_ctx* c, int64_t* n_out) {
  if (!c) return EMC_ERR_ARG;
  const RestartCells* rs = nullptr;
  const std::string e = restart_table(c, &rs);
  if (!e.empty()) return fail(c, EMC_ERR_IO, e);
  // a restarted run continues at restart_step (src/mcmodel.cpp:507-510)
  c->next_step = rs ? (uint64_t)std::max(0, c->dev.cfg.restart_step) : 0;
  if (c->gpu) {
    const int st = gpu_sample_population(c, rs);
    if (st) return st;
    if (n_out) *n_out = c->gpu->n;
    return EMC_OK;
  }
  sample_population(c, rs);
  if (n_out) *n_out = (int64_t)c->hp.size();
  return EMC_OK;
}

int emc_particles_count(emc_ctx* c, int64_t* n_out) {
  if (!c || !n_out) return EMC_ERR_ARG;
  if (c->gpu) {
    int st = settle_n(c);
    if (st) return st;
  }
  *n_out = c->gpu ? c->gpu->n : (int64_t)c->hp.size();
  return EMC_OK;
}

int emc_particles_upload(emc_ctx* c, const emc_particles* h, int64_t n) {
  if (!c || !h || n < 0) return EMC_ERR_ARG;
  HostParticles& hp = c->hp;
  hp.resize((size_t)n);
  for (int64_t s = 0; s < n; ++s) {
    hp.x[s] = h->x[s]; hp.y[s] = h->y[s]; hp.z[s] = h->z[s]; hp.kx[s] = h->kx[s]; hp.ky[s] = h->ky[s];
    hp.kz[s] = h->kz[s]; hp.energy[s] = h->energy[s]; hp.charge[s] = h->charge[s];
    hp.i[s] = h->i[s]; hp.j[s] = h->j[s]; hp.k[s] = h->k[s]; hp.itet[s] = h->itet[s]; hp.isym[s] = h->isym[s];
    hp.ptype[s] = h->ptype[s]; hp.key[s] = h->key[s];
    const Device& d = c->dev;
    if (hp.i[s] >= 0 && (hp.i[s] >= d.cnx || hp.j[s] < 0 || hp.j[s] >= d.cny || hp.k[s] < 0 || hp.k[s] >= d.cnz ||
                         hp.itet[s] < 0 || hp.itet[s] >= c->bb->band.nt || hp.isym[s] < 0 || hp.isym[s] >= 48 ||
                         hp.ptype[s] < 0 || hp.ptype[s] > 1))
      return fail(c, EMC_ERR_ARG, "particle " + std::to_string(s) + " has an out-of-range index");
  }
  if (c->gpu) return gpu_upload_host(c);
  return EMC_OK;
}

int emc_particles_download(emc_ctx* c, emc_particles* h, int64_t cap, int64_t* n_out) {
  if (!c || !h) return EMC_ERR_ARG;
  if (!c->gpu) {
    const HostParticles& hp = c->hp;
    const int64_t n = (int64_t)hp.size();
    if (n_out) *n_out = n;
    if (n > cap) return fail(c, EMC_ERR_CAPACITY, "host buffer too small");
    for (int64_t s = 0; s < n; ++s) {
      h->x[s] = hp.x[s]; h->y[s] = hp.y[s]; h->z[s] = hp.z[s]; h->kx[s] = hp.kx[s]; h->ky[s] = hp.ky[s];
      h->kz[s] = hp.kz[s]; h->energy[s] = hp.energy[s]; h->charge[s] = hp.charge[s];
      h->i[s] = hp.i[s]; h->j[s] = hp.j[s]; h->k[s] = hp.k[s]; h->itet[s] = hp.itet[s]; h->isym[s] = hp.isym[s];
      h->ptype[s] = hp.ptype[s]; h->key[s] = hp.key[s];
    }
    return EMC_OK;
  }
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  const long long n = g.n;
  if (n_out) *n_out = n;
  if (n > cap) return fail(c, EMC_ERR_CAPACITY, "host buffer too small");
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  std::vector<uint32_t> pk(n);
  auto dn = [&](void* hptr, const void* dptr, size_t bytes) { return hipMemcpy(hptr, dptr, bytes, hipMemcpyDeviceToHost); };
  HIPCHECK(c, dn(h->x, g.P.x, n * 8)); HIPCHECK(c, dn(h->y, g.P.y, n * 8)); HIPCHECK(c, dn(h->z, g.P.z, n * 8));
  HIPCHECK(c, dn(h->kx, g.P.kx, n * 8)); HIPCHECK(c, dn(h->ky, g.P.ky, n * 8)); HIPCHECK(c, dn(h->kz, g.P.kz, n * 8));
  HIPCHECK(c, dn(h->energy, g.P.energy, n * 8)); HIPCHECK(c, dn(h->charge, g.P.charge, n * 8));
  HIPCHECK(c, dn(pk.data(), g.P.pk, n * 4)); HIPCHECK(c, dn(h->itet, g.P.itet, n * 4));
  HIPCHECK(c, dn(h->isym, g.P.isym, n * 4)); HIPCHECK(c, dn(h->key, g.P.key, n * 8));
  for (long long s = 0; s < n; ++s) {
    if (pk[s] == PK_DEAD) { h->i[s] = -1; h->j[s] = -1; h->k[s] = -1; h->ptype[s] = -1; continue; }
    h->i[s] = pk_i(pk[s]); h->j[s] = pk_j(pk[s]); h->k[s] = pk_k(pk[s]); h->ptype[s] = pk_type(pk[s]);
  }
  return EMC_OK;
}

int emc_fields_set(emc_ctx* c, const emc_cell_fields* f) {
  if (!c || !f) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t NC = c->dev.ncell;
  if (f->hx != f->ex || f->hy != f->ey || f->hz != f->ez || g.hx) {
    // distinct hole fields: give them their own storage once
    if (g.F.hx == g.ex) {
      g.F.hx = g.mem.alloc<double>(NC); g.F.hy = g.mem.alloc<double>(NC); g.F.hz = g.mem.alloc<double>(NC);
      if (g.mem.last != hipSuccess) return fail(c, EMC_ERR_HIP, "allocation of hole fields");
    }
    HIPCHECK(c, hipMemcpy((void*)g.F.hx, f->hx, NC * 8, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy((void*)g.F.hy, f->hy, NC * 8, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy((void*)g.F.hz, f->hz, NC * 8, hipMemcpyHostToDevice));
  }
  HIPCHECK(c, hipMemcpy(g.ex, f->ex, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.ey, f->ey, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.ez, f->ez, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.rho, f->rho, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.rates, f->e_rough, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.rates + NC, f->e_sph, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.rates + 2 * NC, f->h_rough, NC * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.rates + 3 * NC, f->h_sph, NC * 8, hipMemcpyHostToDevice));
  return EMC_OK;
}

int emc_fields_get(emc_ctx* c, emc_cell_fields* f) {
  if (!c || !f) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t NC = c->dev.ncell;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  HIPCHECK(c, hipMemcpy(f->ex, g.ex, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->ey, g.ey, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->ez, g.ez, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->hx, g.F.hx, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->hy, g.F.hy, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->hz, g.F.hz, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->rho, g.rho, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->e_rough, g.rates, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->e_sph, g.rates + NC, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->h_rough, g.rates + 2 * NC, NC * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(f->h_sph, g.rates + 3 * NC, NC * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- step pieces
// Brings the device copy of the table/field descriptors up to date (they change
// only when hole fields get storage of their own or are aliased back).
static int sync_fly_const(emc_ctx* c) {
  Gpu& g = *c->gpu;
  FlyConst k{};
  k.B = g.B; k.M = g.M; k.F = g.F;
  if (std::memcmp(&k, &g.cst_host, sizeof(k)) == 0) return EMC_OK;
  HIPCHECK(c, hipStreamSynchronize(g.stream));  // no flight kernel may still read the old copy
  HIPCHECK(c, hipMemcpy(g.cst, &k, sizeof(k), hipMemcpyHostToDevice));
  g.cst_host = k;
  return EMC_OK;
}

static FlyArgs fly_args(emc_ctx* c, uint64_t step, int flag_heat) {
  Gpu& g = *c->gpu;
  FlyArgs a{};
  a.C = g.cst; a.T = g.T; a.P = g.P;
  if (!flag_heat) { a.T.heat_e = nullptr; a.T.heat_h = nullptr; }
  a.n = g.n; a.cap = g.cap; a.step = step; a.flag_heat = flag_heat;
  a.clones = g.clones; a.nclones = g.ctr; a.clone_cap = g.clone_cap;
  a.dead = g.dead; a.ndead = g.ctr + 1; a.dead_cap = g.dead_cap;
  a.slow = g.slow; a.nslow = g.ctr + 3;
  a.wave_slots = g.wave_slots;
  return a;
}

// Queues one flight step on the stream: fly, two clone passes and the merge,
// all sized on the device, then pinned copies of the counters and tallies.  No
// host round trip; finish_fly() completes the step once the stream has passed.
static int launch_fly(emc_ctx* c, uint64_t step, int flag_heat, FlyArgs& a) {
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  st = sync_fly_const(c);
  if (st) return st;
  // every caller has cleared the block with the counters (zero_tallies)
  a = fly_args(c, step, flag_heat);
  HIPCHECK(c, hipEventRecord(g.ev[5], g.stream));
  HIPCHECK(c, emc_launch_fly(a, g.stream));
  HIPCHECK(c, hipEventRecord(g.ev[6], g.stream));
  HIPCHECK(c, emc_launch_clones_dev(a, g.stream));
  HIPCHECK(c, hipMemcpyAsync(g.tally_host, g.contact, FLYOUT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             g.stream));
  return EMC_OK;
}

// After the stream has passed launch_fly's copies: clones of a third generation
// (rare) fly in host-driven passes and are merged; g.n takes the appended
// survivors.  Returns the number of clone passes that had clones, or an error.
static int finish_fly(emc_ctx* c, const FlyArgs& a) {
  Gpu& g = *c->gpu;
  unsigned long long* h = g.ctr_host;
  g.last_full_loop = (long long)h[3];
  int passes = (int)h[6];
  if (h[5] > h[4]) {
    const unsigned long long first = h[4];
    unsigned long long begin = h[4], end = h[5];
    while (end > begin) {
      HIPCHECK(c, emc_launch_fly_clones(a, begin, end, g.stream));
      HIPCHECK(c, hipMemcpyAsync(h, g.ctr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, g.stream));
      HIPCHECK(c, hipStreamSynchronize(g.stream));
      begin = end;
      end = std::min(h[0], g.clone_cap);
      if (end > begin && ++passes > 64) return fail(c, EMC_ERR_STATE, "clone passes do not terminate");
    }
    HIPCHECK(c, emc_launch_merge(a, first, end, g.ctr + 1, g.ctr + 2, g.stream));
    HIPCHECK(c, hipMemcpyAsync(g.tally_host, g.contact, FLYOUT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               g.stream));
    HIPCHECK(c, hipStreamSynchronize(g.stream));
  }
  const unsigned long long surv = h[2];
  const unsigned long long used = std::min(h[1], g.dead_cap);
  if (surv > used) g.n = std::min<long long>(g.cap, g.n + (long long)(surv - used));
  return passes;
}

// One flight step, synchronous (the emc_fly_step entry point).
static int run_fly(emc_ctx* c, uint64_t step, int flag_heat) {
  Gpu& g = *c->gpu;
  FlyArgs a{};
  int st = launch_fly(c, step, flag_heat, a);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  return finish_fly(c, a);
}

// the flight's tallies from the pinned copies finish_fly() left
static void read_tallies(emc_ctx* c, emc_fly_tallies* out, int passes) {
  Gpu& g = *c->gpu;
  const unsigned long long* ct = g.tally_host;
  const unsigned long long* cn = g.tally_host + EMC_MAX_CONTACT * EMC_CT_NCOL;
  if (!out) return;
  for (int i = 0; i < EMC_MAX_CONTACT; ++i)
    for (int j = 0; j < EMC_CT_NCOL; ++j) out->contact[i][j] += (int64_t)ct[i * EMC_CT_NCOL + j];
  for (int r = 0; r < EMC_NREASON; ++r) out->drops[r] += (int64_t)cn[CNT_DROP0 + r];
  out->n_catch += (int64_t)cn[CNT_CATCH];
  out->n_gen += (int64_t)cn[CNT_GEN];
  out->n_phonon += (int64_t)cn[CNT_PHONON];
  out->n_impurity += (int64_t)cn[CNT_IMPURITY];
  out->n_events += (int64_t)cn[CNT_EVENTS];
  out->n_clone_passes += passes;
  out->n_full_loop += g.last_full_loop;
}

// Contact columns, counters and the clone/merge counters of the flight: one block.
static int zero_tallies(emc_ctx* c) {
  Gpu& g = *c->gpu;
  HIPCHECK(c, emc_launch_zero(g.contact, FLYOUT_N, g.stream));
  return EMC_OK;
}

// The live count after a tile sort is known on the device first; the host picks it
// up (pinned copy + event) at the next point that needs it.
static int settle_n(emc_ctx* c) {
  Gpu& g = *c->gpu;
  if (!g.n_pending) return EMC_OK;
  HIPCHECK(c, hipEventSynchronize(g.n_ready));
  g.n = (long long)*g.n_host;
  g.sorted_n = g.n;
  g.n_pending = false;
  return EMC_OK;
}

// Counting sort of the population by deposit tile into the second buffer, which
// becomes the current one.  Empty slots are dropped.
// One block walks each tile's segment, so the previous segments are used only
// while they hold at most about a thousand particles on average (the 80x160x48
// FinFET: 450; a 20x40x12 device with 300 k particles: 4 000, which the 2048
// striding blocks of the all-movers path sort faster).
static bool seg_usable(const Gpu& g) { return g.seg_valid && g.sorted_n <= 1024LL * g.ntile; }

static SortArgs sort_args(emc_ctx* c) {
  Gpu& g = *c->gpu;
  SortArgs a{};
  a.P = g.P; a.Q = g.P2; a.n = g.n;
  a.ntx = g.ntx; a.nty = g.nty; a.ntz = g.ntz; a.ntile = g.ntile;
  a.tcount = g.tcount; a.toff = g.toff; a.tcursor = g.tcursor; a.tstay = g.tstay; a.toff_old = g.toff_old;
  a.seg_valid = seg_usable(g) ? 1 : 0;
  return a;
}
// slots outside the last sort's segments (all of them without segments)
static long long sort_nappend(emc_ctx* c) {
  const Gpu& g = *c->gpu;
  return seg_usable(g) ? std::max(0LL, g.n - g.sorted_n) : g.n;
}

static int run_sort(emc_ctx* c) {
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  SortArgs a = sort_args(c);
  HIPCHECK(c, emc_launch_tile_sort(a, sort_nappend(c), g.stream));
  g.seg_valid = true;
  std::swap(g.P, g.P2);
  HIPCHECK(c, hipMemcpyAsync(g.n_host, g.toff + g.ntile, sizeof(unsigned), hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(c, hipEventRecord(g.n_ready, g.stream));
  g.n_pending = true;
  return EMC_OK;
}

static DepositArgs deposit_args(emc_ctx* c, int stat_accumulate) {
  Gpu& g = *c->gpu;
  DepositArgs a{};
  a.B = g.B; a.M = g.M; a.P = g.P;
  a.nnode = c->dev.nnode; a.ncell = c->dev.ncell;
  a.ntx = g.ntx; a.nty = g.nty; a.ntz = g.ntz; a.ntile = g.ntile;
  a.toff = g.toff; a.blk = g.blk;
  a.node_q = g.node_q; a.node_stat = g.node_stat; a.cell_q = g.cell_q; a.cell_n = g.cell_n;
  a.stat_accumulate = stat_accumulate;
  return a;
}

static FieldArgs field_args(emc_ctx* c, const double* pot, bool rho_from_cells) {
  Gpu& g = *c->gpu;
  // without the quantum correction one field serves both species
  if (g.hx) { g.F.hx = g.hx; g.F.hy = g.hy; g.F.hz = g.hz; }
  else { g.F.hx = g.ex; g.F.hy = g.ey; g.F.hz = g.ez; }
  const Device& d = c->dev;
  const Phys& p = d.phys;
  FieldArgs a{};
  a.ncell = d.ncell; a.cnx = d.cnx; a.cny = d.cny; a.cnz = d.cnz; a.pnx = d.pnx; a.pny = d.pny; a.pnz = d.pnz;
  a.pot = pot; a.dx = g.dx; a.dy = g.dy; a.dz = g.dz; a.cvol = g.cvol; a.da = g.da; a.eeff_dir = g.eeff;
  a.cell_q = rho_from_cells ? g.cell_q : nullptr;
  a.ex = g.ex; a.ey = g.ey; a.ez = g.ez; a.rho = g.rho;
  a.rates = d.cfg.flag_surface ? g.rates : nullptr;
  a.flag_rough = d.cfg.flag_rough; a.flag_sphonon = d.cfg.flag_sphonon;
  a.eps_si = p.eps[SILICON]; a.eps_ox = p.eps[OXIDE]; a.Tn = p.Tn; a.T0 = p.T0; a.field0 = p.field0;
  a.em0 = p.em0; a.ec0 = p.ec0; a.hq0 = p.hq0; a.conc0 = p.conc0; a.scrt0 = p.scrt0; a.sirho = p.sirho;
  a.surf_delta = p.SurfSc_delta; a.surf_ail = p.SurfSc_ail; a.surf_XIph = p.SurfSc_XIph;
  a.surf_theta = p.SurfSc_theta; a.surf_gama = p.SurfSc_gama; a.surf_Nbmod = p.SurfSc_Nbmod;
  a.surf_Npf = p.SurfSc_Npf; a.surf_Pft = p.SurfSc_Pft; a.surf_Pfn = p.SurfSc_Pfn; a.surf_Kpha = p.SurfSc_Kpha;
  a.mass_eff[0] = p.mass_eff[0]; a.mass_eff[1] = p.mass_eff[1];
  return a;
}

static PcgArgs pcg_args(emc_ctx* c, const double* b, double* x, double rtol) {
  Gpu& g = *c->gpu;
  const Device& d = c->dev;
  PcgArgs a{};
  a.n = d.nnode; a.pnx = d.pnx; a.pny = d.pny; a.pnz = d.pnz; a.nblocks = g.pcg_blocks;
  a.diag = g.diag; a.cx = g.cx; a.cy = g.cy; a.cz = g.cz; a.b = b; a.kind = g.kind;
  a.x = x; a.r = g.r; a.p = g.p; a.p2 = g.p2; a.q = g.q; a.z = g.z;
  a.part_a = g.part_a; a.part_pq = g.part_pq; a.part_rr = g.part_rr; a.part_rz0 = g.part_rz0; a.part_rz1 = g.part_rz1;
  a.scal = g.scal; a.done = g.done; a.rtol = rtol;
  return a;
}

// Multigrid-preconditioned CG to ||r|| <= rtol ||b|| (the solve of
// PoissonSolver::solve_poisson, src/PoissonSolver.cpp:70-78, on the eliminated
// system).  Iteration k+1 is queued before the host reads the residual of
// iteration k, so the GPU never waits for the check; once converged the device
// flag turns the queued kernels into no-ops.
// With `defer` the solve is only queued past its last check: the result (iteration
// count, residual) is read by finish_pcg() once the stream has passed it.
static int finish_pcg(emc_ctx* c, double rtol, int* iters, double* relres);
template <class PA, class MA>
static int run_pcg_with(emc_ctx* c, const PA& a, const MA& mg, double rtol, int maxit, int* iters, double* relres,
                        bool defer);
static int run_pcg(emc_ctx* c, const double* b, double* x, double rtol, int maxit, int* iters, double* relres,
                   bool defer = false) {
  Gpu& g = *c->gpu;
  if (c->opts.poisson_fp32 == 2) {
    // FP64 CG, FP32 cycle
    const int st = ensure_f32(c);
    if (st) return st;
    PcgArgs a = pcg_args(c, b, x, rtol);
    a.r32 = g.r32;
    a.z32 = g.z32;
    return run_pcg_with(c, a, g.pmg, rtol, maxit, iters, relres, defer);
  }
  return run_pcg_with(c, pcg_args(c, b, x, rtol), g.mg, rtol, maxit, iters, relres, defer);
}

// The solve on explicit work vectors: `a` and `mg` name every buffer the
// kernels write (the solve of the context uses its own, emc_poisson_solve_ws a
// caller's workspace).
template <class PA, class MA>
static int run_pcg_with(emc_ctx* c, const PA& a, const MA& mg, double rtol, int maxit, int* iters, double* relres,
                        bool defer) {
  Gpu& g = *c->gpu;
  // k_pcg_begin sets done when the start already meets the tolerance: the
  // queued iterations are then no-ops and the first check below stops the loop
  HIPCHECK(c, emc_launch_pcg_start(a, mg, g.stream));
  double* sh = g.scal_host;
  // Each check reads the residual of the iteration before the one just queued, so
  // the GPU never waits for the host; the iteration the previous solve converged in
  // is checked as soon as it is queued instead, so that a solve converging there
  // queues no iteration that would only find the done flag set.
  const int guess = c->last_pcg_iters;
  int checked = -1, last = -1;
  for (int k = 0; k < maxit; ++k) {
    last = k;
    HIPCHECK(c, emc_launch_pcg_iter(a, mg, k, g.stream));
    double* slot = sh + 8 * (k & 1);
    HIPCHECK(c, hipMemcpyAsync(slot, a.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHECK(c, hipEventRecord(g.scal_ready[k & 1], g.stream));
    const int chk = (k + 1 == guess) ? k : k - 1;
    if (chk < 0 || chk <= checked) continue;
    HIPCHECK(c, hipEventSynchronize(g.scal_ready[chk & 1]));
    checked = chk;
    const double* res = sh + 8 * (chk & 1);
    if (!std::isfinite(res[0])) return fail(c, EMC_ERR_SOLVER, "PCG residual is not finite");
    if (res[0] <= res[4]) break;
  }
  // the kernels queued after convergence leave scal as it was, so the copy of the
  // last queued iteration holds the result (read once the stream has passed it)
  if (last < 0) HIPCHECK(c, hipMemcpyAsync(sh + 16, a.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
  g.pcg_result = last < 0 ? sh + 16 : sh + 8 * (last & 1);
  g.pcg_pending = true;
  if (defer) return EMC_OK;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  return finish_pcg(c, rtol, iters, relres);
}

static int finish_pcg(emc_ctx* c, double rtol, int* iters, double* relres) {
  Gpu& g = *c->gpu;
  if (!g.pcg_pending) return EMC_OK;
  g.pcg_pending = false;
  const double* sh = g.pcg_result;
  const double bn = std::sqrt(sh[2]) > 0 ? std::sqrt(sh[2]) : 1.0;
  const int it = (int)sh[3];
  const double rel = std::sqrt(std::max(0.0, sh[0])) / bn;
  if (!std::isfinite(rel)) return fail(c, EMC_ERR_SOLVER, "PCG residual is not finite");
  if (iters) *iters = it;
  if (relres) *relres = rel;
  c->last_pcg_iters = it;
  c->last_relres = rel;
  if (rel > rtol) return fail(c, EMC_ERR_SOLVER, "PCG did not reach the tolerance");
  return EMC_OK;
}

// The FP32 copies of the mixed-precision solve (first use).
static int ensure_f32(emc_ctx* c) {
  Gpu& g = *c->gpu;
  if (g.f32_ready) return EMC_OK;
  const Device& d = c->dev;
  const size_t NN = d.nnode;
  DevBuf& m = g.mem;
  auto f32 = [&](const std::vector<double>& v) {
    std::vector<float> f(v.size());
    for (size_t q = 0; q < v.size(); ++q) f[q] = (float)v[q];
    return m.dup(f);
  };
  PcgArgsT<float>& a = g.fa;
  a.n = NN; a.pnx = d.pnx; a.pny = d.pny; a.pnz = d.pnz; a.nblocks = g.pcg_blocks;
  a.diag = f32(d.diag); a.cx = f32(d.cx); a.cy = f32(d.cy); a.cz = f32(d.cz);
  a.kind = g.kind;
  float* b = m.alloc<float>(NN);
  a.b = b;
  a.x = m.alloc<float>(NN); a.r = m.alloc<float>(NN); a.p = m.alloc<float>(NN); a.p2 = m.alloc<float>(NN);
  a.q = m.alloc<float>(NN); a.z = m.alloc<float>(NN);
  for (double** pp : {&a.part_a, &a.part_pq, &a.part_rr, &a.part_rz0, &a.part_rz1}) *pp = m.alloc<double>(3 * g.pcg_blocks);
  a.scal = m.alloc<double>(8);
  a.done = m.alloc<int>(1);
  const MgHierarchy& h = c->mg;
  MgArgsT<float>& F = g.fmg;
  F.nl = g.mg.nl; F.nc = g.mg.nc;
  F.cinv = f32(h.cinv);
  F.diag = a.diag; F.cx = a.cx; F.cy = a.cy; F.cz = a.cz;
  F.done = a.done;
  for (int l = 0; l < F.nl; ++l) {
    const MgLevelDev& D = g.mg.L[l];
    MgLevelDevT<float>& E = F.L[l];
    E.nx = D.nx; E.ny = D.ny; E.nz = D.nz; E.n = D.n;
    E.c14 = D.c14;
    E.wdinv = f32(h.lv[l].wdinv);
    E.x = m.alloc<float>(D.n); E.t = m.alloc<float>(D.n); E.r = m.alloc<float>(D.n);
    E.b = l > 0 ? m.alloc<float>(D.n) : a.r;
  }
  g.r32 = m.alloc<float>(NN);
  g.z32 = m.alloc<float>(NN);
  g.pmg = F;
  g.pmg.L[0].b = g.r32;
  g.pmg.done = g.done;
  if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("FP32 solver buffers: ") + hipGetErrorString(m.last));
  g.f32_ready = true;
  return EMC_OK;
}

// Mixed precision (the FP32 bulk with FP64 iterative refinement): r = b - A x in
// FP64; while ||r|| > rtol ||b||, the multigrid-preconditioned CG solves A d = r
// with FP32 vectors and operator to inner_rtol, and x += d in FP64.  Each outer
// step gains about inner_rtol in the FP64 residual, so the result meets the FP64
// tolerance; FP32 halves the bytes each CG iteration streams.
static int run_pcg_mixed(emc_ctx* c, const double* b, double* x, double rtol, int maxit_outer, double inner_rtol,
                         int* outer_out, int* inner_out, double* relres) {
  Gpu& g = *c->gpu;
  int st = ensure_f32(c);
  if (st) return st;
  RefineArgs ra{};
  ra.a = pcg_args(c, b, x, rtol);
  ra.r32 = const_cast<float*>(g.fa.b);
  ra.d32 = g.fa.x;
  PcgArgsT<float> fa = g.fa;
  fa.rtol = inner_rtol;
  double* sh = g.scal_host + 24;  // [24, 32): the refinement's norms
  int outer = 0, inner = 0;
  double rel = 0;
  for (;;) {
    HIPCHECK(c, emc_launch_refine_resid(ra, g.stream));
    HIPCHECK(c, hipMemcpyAsync(sh, ra.a.scal, 8 * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHECK(c, hipStreamSynchronize(g.stream));
    const double bn = std::sqrt(sh[2]) > 0 ? std::sqrt(sh[2]) : 1.0;
    rel = std::sqrt(std::max(0.0, sh[0])) / bn;
    if (!std::isfinite(rel)) return fail(c, EMC_ERR_SOLVER, "residual of the refinement is not finite");
    if (rel <= rtol) break;
    if (outer >= maxit_outer) return fail(c, EMC_ERR_SOLVER, "iterative refinement did not reach the tolerance");
    HIPCHECK(c, emc_launch_zero_f32(fa.x, fa.n, g.stream));
    int it = 0;
    st = run_pcg_with(c, fa, g.fmg, inner_rtol, c->opts.poisson_maxit, &it, nullptr, false);
    if (st) return st;
    inner += it;
    HIPCHECK(c, emc_launch_refine_add(ra, g.stream));
    ++outer;
  }
  if (outer_out) *outer_out = outer;
  if (inner_out) *inner_out = inner;
  if (relres) *relres = rel;
  c->last_relres = rel;
  return EMC_OK;
}

// MultipleRefresh (src/MultipleRefresh.cpp:61-143) of the device population:
// resampled into the second particle buffer, which becomes the current one.
static int run_mr(emc_ctx* c, uint64_t step, long long* before, long long* after) {
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  const Device& d = c->dev;
  MrArgs& a = g.mr;
  if (g.mr_ready && g.mr_cap < g.cap) {
    // the population grew: slot-sized work arrays of the new capacity
    DevBuf& m = g.mem;
    for (const void* p : {(const void*)a.gk, (const void*)a.gk2, (const void*)a.idx, (const void*)a.idx2,
                          (const void*)a.sk, (const void*)a.sk2, (const void*)a.temp})
      m.release(p);
    const long long cap = g.cap, ng = 2 * (long long)d.ncell;
    a.gk = m.alloc<unsigned>(cap); a.gk2 = m.alloc<unsigned>(cap);
    a.idx = m.alloc<unsigned>(cap); a.idx2 = m.alloc<unsigned>(cap);
    a.sk = m.alloc<unsigned long long>(cap); a.sk2 = m.alloc<unsigned long long>(cap);
    a.temp_bytes = emc_mr_temp_bytes(cap, ng);
    a.temp = m.alloc<unsigned char>(a.temp_bytes > 0 ? a.temp_bytes : 1);
    if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("MultipleRefresh buffers: ") + hipGetErrorString(m.last));
    g.mr_cap = cap;
  }
  if (!g.mr_ready) {
    DevBuf& m = g.mem;
    const long long cap = g.cap, ng = 2 * (long long)d.ncell;
    a.gk = m.alloc<unsigned>(cap); a.gk2 = m.alloc<unsigned>(cap);
    a.idx = m.alloc<unsigned>(cap); a.idx2 = m.alloc<unsigned>(cap);
    a.sk = m.alloc<unsigned long long>(cap); a.sk2 = m.alloc<unsigned long long>(cap);
    g.mr_cap = cap;
    a.gstart = m.alloc<unsigned>(ng); a.gend = m.alloc<unsigned>(ng);
    a.cnt = m.alloc<unsigned long long>(ng); a.off = m.alloc<unsigned long long>(ng);
    a.tot = m.alloc<double>(ng); a.keep = m.alloc<int>(ng);
    a.nlive = m.alloc<unsigned long long>(1);
    a.desired_e = m.dup(c->mr_desired_e);
    a.desired_h = m.dup(c->mr_desired_h);
    a.temp_bytes = emc_mr_temp_bytes(cap, ng);
    a.temp = m.alloc<unsigned char>(a.temp_bytes > 0 ? a.temp_bytes : 1);
    if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("MultipleRefresh buffers: ") + hipGetErrorString(m.last));
    g.mr_ready = true;
  }
  a.P = g.P; a.Q = g.P2; a.n = g.n;
  a.ngroup = 2 * (long long)d.ncell; a.cnx = d.cnx; a.cnz = d.cnz; a.step = step;
  if (a.n <= 0) {
    if (before) *before = 0;
    if (after) *after = 0;
    return EMC_OK;
  }
  HIPCHECK(c, emc_launch_mr_plan(a, g.stream));
  unsigned long long h[3];
  HIPCHECK(c, hipMemcpyAsync(&h[0], a.nlive, 8, hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(c, hipMemcpyAsync(&h[1], a.off + (a.ngroup - 1), 8, hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(c, hipMemcpyAsync(&h[2], a.cnt + (a.ngroup - 1), 8, hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  const long long total = (long long)(h[1] + h[2]);
  if (total > g.cap) {
    // The refreshed population needs more slots than it was given: it is written
    // into new buffers of 1.25x its size, and the flight's work lists grow with them.
    const long long cap = (long long)std::ceil(total * 1.25) + 1024;
    if (cap > 0x7fffffffLL) return fail(c, EMC_ERR_CAPACITY, "MultipleRefresh: population above 2^31-1 slots");
    DevBuf& m = g.mem;
    PartDev out{}, spare{};
    alloc_part(m, out, cap);
    alloc_part(m, spare, cap);
    const unsigned long long clone_cap = (unsigned long long)std::max<long long>(4096, cap / 16);
    CloneRec* clones = m.alloc<CloneRec>(clone_cap);
    long long* dead = m.alloc<long long>(clone_cap);
    uint32_t* slow = m.alloc<uint32_t>(cap);
    if (m.last != hipSuccess) return fail(c, EMC_ERR_HIP, std::string("particle allocation: ") + hipGetErrorString(m.last));
    a.Q = out;
    HIPCHECK(c, emc_launch_mr_write(a, g.stream));
    HIPCHECK(c, hipStreamSynchronize(g.stream));
    release_part(m, g.P);
    release_part(m, g.P2);
    m.release(g.clones); m.release(g.dead); m.release(g.slow);
    g.P = out;
    g.P2 = spare;
    g.cap = cap;
    g.clones = clones; g.clone_cap = clone_cap;
    g.dead = dead; g.dead_cap = clone_cap;
    g.slow = slow;
    g.n = total;
    g.seg_valid = false;
    if (before) *before = (long long)h[0];
    if (after) *after = total;
    return EMC_OK;
  }
  HIPCHECK(c, emc_launch_mr_write(a, g.stream));
  std::swap(g.P, g.P2);
  g.n = total;
  g.seg_valid = false;
  if (before) *before = (long long)h[0];
  if (after) *after = total;
  return EMC_OK;
}

extern "C" {

// ---- checkpoint: the device state a later load continues from exactly
static const char CKPT_MAGIC[8] = {'E', 'M', 'C', 'C', 'K', 'P', 'T', '1'};
static const uint32_t CKPT_VERSION = 4;

// FNV-1a over what a resumed run must share with the saved one: the mesh, the
// doping and materials, the contacts and their voltages, the time step, the
// temperature, the refresh, statistics and physics switches and the run seed
// (total_step and the output cadence of the heating phase may change).
static uint64_t deck_fingerprint(const emc_ctx* c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t q = 0; q < n; ++q) { h ^= b[q]; h *= 1099511628211ull; }
  };
  const Device& d = c->dev;
  const Config& f = d.cfg;
  for (const auto* v : {&d.lx, &d.ly, &d.lz, &d.donor, &d.acceptor}) mix(v->data(), v->size() * sizeof(double));
  mix(d.material.data(), d.material.size() * sizeof(int));
  mix(d.motion.data(), d.motion.size());
  for (const ContactDef& k : d.contacts) {
    mix(&k.phims, sizeof(double));
    mix(&k.vapp, sizeof(double));
    mix(&k.nplane, sizeof(int));
  }
  const double sc[2] = {f.dt, f.temperature};
  mix(sc, sizeof(sc));
  const int fl[9] = {f.flag_mr, f.mr_step, f.flag_surface, f.flag_rough, f.flag_sphonon, f.flag_simp, f.flag_qc,
                     f.stat_step, f.quantum_start_step};
  mix(fl, sizeof(fl));
  mix(&c->opts.seed, sizeof(c->opts.seed));
  return h;
}

int emc_checkpoint_save(emc_ctx* c, const char* path) {
  if (!c || !path) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the checkpoint holds device state");
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  const Device& d = c->dev;
  const long long n = g.n;
  const size_t NN = d.nnode;
  std::vector<uint32_t> pk(n);
  HIPCHECK(c, hipMemcpy(pk.data(), g.P.pk, n * 4, hipMemcpyDeviceToHost));
  std::vector<long long> live;
  live.reserve(n);
  for (long long s = 0; s < n; ++s)
    if (pk[s] != PK_DEAD) live.push_back(s);
  const long long nl = (long long)live.size();
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, EMC_ERR_IO, std::string("cannot write ") + path);
  bool ok = true;
  auto put = [&](const void* p, size_t bytes) { ok = ok && std::fwrite(p, 1, bytes, f) == bytes; };
  const uint32_t version = CKPT_VERSION;
  const int32_t dims[4] = {d.cnx, d.cny, d.cnz, 0};
  const uint64_t seed = c->opts.seed, next = c->next_step;
  const uint64_t fp = deck_fingerprint(c);
  put(CKPT_MAGIC, 8); put(&version, 4); put(dims, 16); put(&seed, 8); put(&next, 8); put(&nl, 8); put(&fp, 8);
  // live particles, compacted in slot order
  auto put_live = [&](const void* dptr, size_t width) {
    std::vector<unsigned char> all((size_t)n * width), sel((size_t)nl * width);
    if (hipMemcpy(all.data(), dptr, (size_t)n * width, hipMemcpyDeviceToHost) != hipSuccess) { ok = false; return; }
    for (long long q = 0; q < nl; ++q) std::memcpy(&sel[(size_t)q * width], &all[(size_t)live[q] * width], width);
    put(sel.data(), sel.size());
  };
  for (const void* a : {(const void*)g.P.x, (const void*)g.P.y, (const void*)g.P.z, (const void*)g.P.kx,
                        (const void*)g.P.ky, (const void*)g.P.kz, (const void*)g.P.energy, (const void*)g.P.charge})
    put_live(a, 8);
  put_live(g.P.pk, 4); put_live(g.P.itet, 4); put_live(g.P.isym, 4); put_live(g.P.key, 8);
  auto put_dev = [&](const void* dptr, size_t bytes) {
    std::vector<unsigned char> h(bytes);
    if (hipMemcpy(h.data(), dptr, bytes, hipMemcpyDeviceToHost) != hipSuccess) { ok = false; return; }
    put(h.data(), bytes);
  };
  put_dev(g.pot, NN * 8);
  put_dev(g.node_stat, 10 * NN * 8);
  put_dev(g.heat_e, NN * 8); put_dev(g.heat_h, NN * 8);
  put_dev(g.rcur, (size_t)d.pny * 8); put_dev(g.lcur, (size_t)d.pny * 8);
  // the statistics window, what output_stat keeps, and the heat phase
  put_dev(g.stat_pot, NN * 8);
  put_dev(g.stat_enum, d.ncell * 8);
  put_dev(g.stat_heat_e, NN * 8); put_dev(g.stat_heat_h, NN * 8);
  put(&c->win, sizeof(c->win));
  const int64_t flags[4] = {c->win_steps, g.heat_live ? 1 : 0, c->heat_phase ? 1 : 0, (int64_t)c->heat_next};
  put(flags, sizeof(flags));
  for (const std::vector<double>* v : {&c->saved.pot, &c->saved.e_charge, &c->saved.h_charge}) {
    std::vector<double> h(*v);
    h.resize(NN, 0.0);
    put(h.data(), NN * 8);
  }
  if (c->heat_phase) put_dev(g.pot_avg, NN * 8);
  // quantum correction: the corrected potential, its window sum, its window mean
  const int64_t has_qc = g.qc_pot ? 1 : 0;
  put(&has_qc, 8);
  if (has_qc) {
    put_dev(g.qc_pot, NN * 8);
    put_dev(g.stat_qc_pot, NN * 8);
    std::vector<double> h(c->saved.qc_pot);
    h.resize(NN, 0.0);
    put(h.data(), NN * 8);
    if (c->heat_phase) put_dev(g.qc_avg, NN * 8);
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? EMC_OK : fail(c, EMC_ERR_IO, std::string("writing ") + path + " failed");
}

int emc_checkpoint_load(emc_ctx* c, const char* path) {
  if (!c || !path) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the checkpoint holds device state");
  Gpu& g = *c->gpu;
  int st = settle_n(c);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  const Device& d = c->dev;
  const size_t NN = d.nnode;
  // The whole file is read and checked first; the context changes only once all
  // of it is there (a truncated or foreign file leaves the context as it was).
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, EMC_ERR_IO, std::string("cannot read ") + path);
  bool ok = true;
  auto get = [&](void* p, size_t bytes) { ok = ok && std::fread(p, 1, bytes, f) == bytes; };
  char magic[8];
  uint32_t version = 0;
  int32_t dims[4];
  uint64_t seed = 0, next = 0, fp = 0;
  long long nl = 0;
  get(magic, 8); get(&version, 4); get(dims, 16); get(&seed, 8); get(&next, 8); get(&nl, 8); get(&fp, 8);
  if (!ok || std::memcmp(magic, CKPT_MAGIC, 8) != 0 || version != CKPT_VERSION) {
    std::fclose(f);
    return fail(c, EMC_ERR_IO, std::string(path) + " is not a checkpoint of this format");
  }
  if (dims[0] != d.cnx || dims[1] != d.cny || dims[2] != d.cnz || nl < 0 || nl > 0x7fffffffLL) {
    std::fclose(f);
    return fail(c, EMC_ERR_DECK, "checkpoint mesh does not match the deck");
  }
  if (fp != deck_fingerprint(c)) {
    std::fclose(f);
    return fail(c, EMC_ERR_DECK, "checkpoint was saved under another deck or seed (mesh, doping, contacts, dt, "
                                 "temperature, switches)");
  }
  // device destination (null: host only) and the bytes read for it, in file order
  struct Part {
    void** dev;
    size_t bytes;
    std::vector<unsigned char> h;
  };
  std::vector<Part> parts;
  double* pot_avg_tmp = nullptr;
  double* qc_avg_tmp = nullptr;
  auto take = [&](void** dev, size_t bytes) {
    Part p{dev, bytes, std::vector<unsigned char>(bytes)};
    get(p.h.data(), bytes);
    parts.push_back(std::move(p));
  };
  PartDev Pn = g.P;  // pointers bound below, once the capacity is settled
  void* slots[12];
  const size_t w8 = (size_t)nl * 8, w4 = (size_t)nl * 4;
  for (int q = 0; q < 8; ++q) take(&slots[q], w8);
  for (int q = 8; q < 11; ++q) take(&slots[q], w4);
  take(&slots[11], w8);
  void* dv[12];
  dv[0] = g.pot; dv[1] = g.node_stat; dv[2] = g.heat_e; dv[3] = g.heat_h; dv[4] = g.rcur; dv[5] = g.lcur;
  dv[6] = g.stat_pot; dv[7] = g.stat_enum; dv[8] = g.stat_heat_e; dv[9] = g.stat_heat_h;
  take(&dv[0], NN * 8);
  take(&dv[1], 10 * NN * 8);
  take(&dv[2], NN * 8); take(&dv[3], NN * 8);
  take(&dv[4], (size_t)d.pny * 8); take(&dv[5], (size_t)d.pny * 8);
  take(&dv[6], NN * 8);
  take(&dv[7], d.ncell * 8);
  take(&dv[8], NN * 8); take(&dv[9], NN * 8);
  emc_fly_tallies win;
  int64_t flags[4] = {0, 0, 0, 0};
  get(&win, sizeof(win));
  get(flags, sizeof(flags));
  SavedStat saved;
  for (std::vector<double>* v : {&saved.pot, &saved.e_charge, &saved.h_charge}) {
    v->resize(NN);
    get(v->data(), NN * 8);
  }
  void* avg_dst = nullptr;
  if (ok && flags[2]) take(&avg_dst, NN * 8);
  int64_t has_qc = 0;
  get(&has_qc, 8);
  if (ok && has_qc != (g.qc_pot ? 1 : 0)) {
    std::fclose(f);
    return fail(c, EMC_ERR_DECK, "checkpoint and deck disagree on Flag_QuantumCorrection");
  }
  void* qdv[3] = {g.qc_pot, g.stat_qc_pot, nullptr};
  if (ok && has_qc) {
    take(&qdv[0], NN * 8);
    take(&qdv[1], NN * 8);
    saved.qc_pot.resize(NN);
    get(saved.qc_pot.data(), NN * 8);
    if (ok && flags[2]) take(&qdv[2], NN * 8);
  }
  const bool at_end = ok && std::fgetc(f) == EOF;
  std::fclose(f);
  if (!ok || !at_end) return fail(c, EMC_ERR_IO, std::string(path) + " is truncated or longer than its header says");
  // the file is whole: now the device
  const long long want = (long long)std::ceil(nl * std::max(1.0, c->opts.capacity_factor)) + 1024;
  st = gpu_alloc_particles(c, want);
  if (st) return st;
  if (nl > g.cap) return fail(c, EMC_ERR_CAPACITY, "checkpoint population exceeds the device capacity");
  Pn = g.P;
  void* pdst[12] = {Pn.x, Pn.y, Pn.z, Pn.kx, Pn.ky, Pn.kz, Pn.energy, Pn.charge, Pn.pk, Pn.itet, Pn.isym, Pn.key};
  for (int q = 0; q < 12; ++q) slots[q] = pdst[q];
  if (flags[2]) {
    if (!g.pot_avg) g.pot_avg = pot_avg_tmp = g.mem.alloc<double>(NN);
    avg_dst = g.pot_avg;
    if (has_qc) {
      if (!g.qc_avg) g.qc_avg = qc_avg_tmp = g.mem.alloc<double>(NN);
      qdv[2] = g.qc_avg;
    }
    if (!g.pot_avg || (has_qc && !g.qc_avg)) return fail(c, EMC_ERR_HIP, "heat-phase buffers");
  }
  (void)pot_avg_tmp;
  (void)qc_avg_tmp;
  for (const Part& p : parts)
    if (p.bytes) HIPCHECK(c, hipMemcpy(*p.dev, p.h.data(), p.bytes, hipMemcpyHostToDevice));
  g.n = nl;
  g.n_pending = false;
  g.seg_valid = false;
  c->next_step = next;
  c->win = win;
  c->win_steps = flags[0];
  g.heat_live = flags[1] != 0;
  c->heat_phase = flags[2] != 0;
  c->heat_next = (uint64_t)flags[3];
  c->saved = std::move(saved);
  c->hp.resize(0);
  return EMC_OK;
}

int emc_multiple_refresh(emc_ctx* c, uint64_t step, int64_t* n_before, int64_t* n_after) {
  if (!c) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: MultipleRefresh runs on the GPU only");
  if (c->gpu->cap == 0) return fail(c, EMC_ERR_STATE, "no particles on the device");
  long long b = 0, a = 0;
  const int st = run_mr(c, step, &b, &a);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(c->gpu->stream));
  if (n_before) *n_before = b;
  if (n_after) *n_after = a;
  return EMC_OK;
}

int emc_fly_step(emc_ctx* c, uint64_t step, int flag_heat, emc_fly_tallies* out) {
  if (!c) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the flight runs on the GPU only");
  if (c->gpu->cap == 0) return fail(c, EMC_ERR_STATE, "no particles on the device");
  if (out) std::memset(out, 0, sizeof(*out));
  int st = zero_tallies(c);
  if (st) return st;
  int passes = run_fly(c, step, flag_heat);
  if (passes < 0) return passes;
  if (flag_heat) c->gpu->heat_live = true;
  read_tallies(c, out, passes);
  HIPCHECK(c, hipGetLastError());
  return EMC_OK;
}

int emc_heat_get(emc_ctx* c, int64_t* heat_e, int64_t* heat_h) {
  if (!c || !c->gpu) return EMC_ERR_ARG;
  Gpu& g = *c->gpu;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  if (heat_e) HIPCHECK(c, hipMemcpy(heat_e, g.heat_e, c->dev.nnode * 8, hipMemcpyDeviceToHost));
  if (heat_h) HIPCHECK(c, hipMemcpy(heat_h, g.heat_h, c->dev.nnode * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

int emc_currents_get(emc_ctx* c, int64_t* rcur, int64_t* lcur) {
  if (!c || !c->gpu) return EMC_ERR_ARG;
  Gpu& g = *c->gpu;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  if (rcur) HIPCHECK(c, hipMemcpy(rcur, g.rcur, c->dev.pny * 8, hipMemcpyDeviceToHost));
  if (lcur) HIPCHECK(c, hipMemcpy(lcur, g.lcur, c->dev.pny * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

int emc_tallies_reset(emc_ctx* c) {
  if (!c || !c->gpu) return EMC_ERR_ARG;
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  HIPCHECK(c, emc_launch_zero(g.heat_e, NN, g.stream));
  HIPCHECK(c, emc_launch_zero(g.heat_h, NN, g.stream));
  HIPCHECK(c, emc_launch_zero(g.rcur, c->dev.pny, g.stream));
  HIPCHECK(c, emc_launch_zero(g.lcur, c->dev.pny, g.stream));
  HIPCHECK(c, emc_launch_zero(g.node_stat, 10 * NN, g.stream));
  HIPCHECK(c, emc_launch_zero((unsigned long long*)g.stat_pot, NN, g.stream));
  HIPCHECK(c, emc_launch_zero(g.stat_enum, c->dev.ncell, g.stream));
  HIPCHECK(c, emc_launch_zero(g.stat_heat_e, NN, g.stream));
  HIPCHECK(c, emc_launch_zero(g.stat_heat_h, NN, g.stream));
  if (g.stat_qc_pot) HIPCHECK(c, emc_launch_zero((unsigned long long*)g.stat_qc_pot, NN, g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  g.heat_live = false;
  std::memset(&c->win, 0, sizeof(c->win));
  c->win_steps = 0;
  return EMC_OK;
}

int emc_stat_window_get(emc_ctx* c, emc_stat_window* w) {
  if (!c || !w) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the window is kept on the device");
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode, NC = c->dev.ncell;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  if (w->stat_pot) HIPCHECK(c, hipMemcpy(w->stat_pot, g.stat_pot, NN * 8, hipMemcpyDeviceToHost));
  if (w->stat_qc_pot)
    HIPCHECK(c, hipMemcpy(w->stat_qc_pot, g.stat_qc_pot ? g.stat_qc_pot : g.stat_pot, NN * 8, hipMemcpyDeviceToHost));
  if (w->node_stat) HIPCHECK(c, hipMemcpy(w->node_stat, g.node_stat, 10 * NN * 8, hipMemcpyDeviceToHost));
  if (w->cell_count) HIPCHECK(c, hipMemcpy(w->cell_count, g.stat_enum, NC * 8, hipMemcpyDeviceToHost));
  if (w->heat_e) HIPCHECK(c, hipMemcpy(w->heat_e, g.stat_heat_e, NN * 8, hipMemcpyDeviceToHost));
  if (w->heat_h) HIPCHECK(c, hipMemcpy(w->heat_h, g.stat_heat_h, NN * 8, hipMemcpyDeviceToHost));
  if (w->rcur) HIPCHECK(c, hipMemcpy(w->rcur, g.rcur, (size_t)c->dev.pny * 8, hipMemcpyDeviceToHost));
  if (w->lcur) HIPCHECK(c, hipMemcpy(w->lcur, g.lcur, (size_t)c->dev.pny * 8, hipMemcpyDeviceToHost));
  w->steps = c->win_steps;
  w->tallies = c->win;
  return EMC_OK;
}

static std::string out_dir(const char* d) { return d && *d ? std::string(d) : std::string("data"); }

int emc_write_stat(emc_ctx* c, const emc_stat_window* w, int heat_file, const char* data_dir, int step) {
  if (!c || !w) return EMC_ERR_ARG;
  if (c->dev.cfg.stat_step <= 0) return fail(c, EMC_ERR_DECK, "stat_step must be positive");
  StatWindow v;
  v.stat_pot = w->stat_pot; v.stat_qc_pot = w->stat_qc_pot; v.node_stat = w->node_stat; v.cell_count = w->cell_count;
  v.heat_e = w->heat_e; v.heat_h = w->heat_h; v.rcur = w->rcur; v.lcur = w->lcur;
  v.tallies = &w->tallies;
  const std::string e = write_stat(c->dev, v, heat_file != 0, step, out_dir(data_dir), c->saved);
  return e.empty() ? EMC_OK : fail(c, EMC_ERR_IO, e);
}

int emc_output_stat(emc_ctx* c, const char* data_dir, int step) {
  if (!c) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the window is kept on the device");
  const size_t NN = c->dev.nnode, NC = c->dev.ncell, NY = c->dev.pny;
  std::vector<double> pot(NN), qpot(NN);
  std::vector<int64_t> ns(10 * NN), cnt(NC), he(NN), hh(NN), rc(NY), lc(NY);
  emc_stat_window w;
  std::memset(&w, 0, sizeof(w));
  w.stat_pot = pot.data(); w.stat_qc_pot = qpot.data(); w.node_stat = ns.data(); w.cell_count = cnt.data();
  w.heat_e = he.data(); w.heat_h = hh.data(); w.rcur = rc.data(); w.lcur = lc.data();
  int st = emc_stat_window_get(c, &w);
  if (st) return st;
  return emc_write_stat(c, &w, c->heat_phase ? 1 : 0, data_dir, step);
}

int emc_write_restart(emc_ctx* c, const emc_particles* h, int64_t n, const char* data_dir, int step) {
  if (!c || !h || n < 0) return EMC_ERR_ARG;
  for (int64_t s = 0; s < n; ++s)
    if (h->i[s] >= 0 && (h->i[s] >= c->dev.cnx || h->j[s] < 0 || h->j[s] >= c->dev.cny || h->k[s] < 0 ||
                         h->k[s] >= c->dev.cnz))
      return fail(c, EMC_ERR_ARG, "particle outside the mesh");
  const std::string e = write_restart(c->dev, *h, n, step, out_dir(data_dir));
  return e.empty() ? EMC_OK : fail(c, EMC_ERR_IO, e);
}

int emc_output_for_restart(emc_ctx* c, const char* data_dir, int step) {
  if (!c) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the population is on the device");
  int64_t n = 0;
  int st = emc_particles_count(c, &n);
  if (st) return st;
  std::vector<double> d8[8];
  std::vector<int32_t> d4[6];
  std::vector<uint64_t> key((size_t)n + 1);
  for (auto& v : d8) v.resize((size_t)n + 1);
  for (auto& v : d4) v.resize((size_t)n + 1);
  emc_particles h{d8[0].data(), d8[1].data(), d8[2].data(), d8[3].data(), d8[4].data(), d8[5].data(),
                  d8[6].data(), d8[7].data(), d4[0].data(), d4[1].data(), d4[2].data(), d4[3].data(),
                  d4[4].data(), d4[5].data(), key.data()};
  int64_t got = 0;
  st = emc_particles_download(c, &h, n, &got);
  if (st) return st;
  return emc_write_restart(c, &h, got, data_dir, step);
}

int emc_output_pvolume(emc_ctx* c, const char* data_dir) {
  if (!c) return EMC_ERR_ARG;
  const std::string e = write_pvolume(c->dev, out_dir(data_dir));
  return e.empty() ? EMC_OK : fail(c, EMC_ERR_IO, e);
}

// heat_field (src/heat.cpp:100-103): the field of the last window's mean potential
int emc_heat_begin(emc_ctx* c) {
  if (!c) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the heat phase runs on the GPU only");
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  if (!g.pot_avg) {
    g.pot_avg = g.mem.alloc<double>(NN);
    if (!g.pot_avg) return fail(c, EMC_ERR_HIP, "device allocation of the heat-phase potential");
  }
  std::vector<double> avg = c->saved.pot;
  avg.resize(NN, 0.0);
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  HIPCHECK(c, hipMemcpy(g.pot_avg, avg.data(), NN * 8, hipMemcpyHostToDevice));
  if (g.qc_pot) {
    if (!g.qc_avg) {
      g.qc_avg = g.mem.alloc<double>(NN);
      if (!g.qc_avg) return fail(c, EMC_ERR_HIP, "device allocation of the heat-phase potential");
    }
    std::vector<double> qa = c->saved.qc_pot.empty() ? c->saved.pot : c->saved.qc_pot;
    qa.resize(NN, 0.0);
    HIPCHECK(c, hipMemcpy(g.qc_avg, qa.data(), NN * 8, hipMemcpyHostToDevice));
  }
  HIPCHECK(c, emc_launch_zero(g.heat_e, NN, g.stream));
  HIPCHECK(c, emc_launch_zero(g.heat_h, NN, g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  c->heat_phase = true;
  c->heat_next = 0;
  return EMC_OK;
}

int emc_heat_step(emc_ctx* c, int nsteps, emc_fly_tallies* acc) {
  if (!c || nsteps < 0) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the heat phase runs on the GPU only");
  if (!c->heat_phase) return fail(c, EMC_ERR_STATE, "emc_heat_begin has not been called");
  if (c->gpu->cap == 0) return fail(c, EMC_ERR_STATE, "no particles on the device");
  return run_steps(c, nsteps, acc, true);
}

int emc_deposit(emc_ctx* c, int64_t* node_q, int64_t* node_stat, int64_t* cell_q, int64_t* cell_n) {
  if (!c) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode, NC = c->dev.ncell;
  int st = run_sort(c);
  if (st) return st;
  HIPCHECK(c, emc_launch_deposit(deposit_args(c, 0), g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  st = settle_n(c);
  if (st) return st;
  if (node_q) HIPCHECK(c, hipMemcpy(node_q, g.node_q, NN * 8, hipMemcpyDeviceToHost));
  if (node_stat) HIPCHECK(c, hipMemcpy(node_stat, g.node_stat, 10 * NN * 8, hipMemcpyDeviceToHost));
  if (cell_q) HIPCHECK(c, hipMemcpy(cell_q, g.cell_q, 3 * NC * 8, hipMemcpyDeviceToHost));
  if (cell_n) HIPCHECK(c, hipMemcpy(cell_n, g.cell_n, NC * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

int emc_rhs_from_charge(emc_ctx* c, const int64_t* node_q, double* rhs_elim) {
  if (!c || !node_q || !rhs_elim) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  HIPCHECK(c, hipMemcpy(g.node_q, node_q, NN * 8, hipMemcpyHostToDevice));
  RhsArgs a{NN, g.node_q, g.pvol, g.pdop, g.dirichlet, g.corr, g.kind, g.rhs, nullptr};
  HIPCHECK(c, emc_launch_rhs(a, g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  HIPCHECK(c, hipMemcpy(rhs_elim, g.rhs, NN * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

int emc_poisson_solve(emc_ctx* c, const double* rhs, double* sol, double rtol, int maxit, int* iters,
                      double* relres) {
  if (!c || !rhs || !sol) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  HIPCHECK(c, hipMemcpy(g.rhs, rhs, NN * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.pot, sol, NN * 8, hipMemcpyHostToDevice));
  int st = run_pcg(c, g.rhs, g.pot, rtol > 0 ? rtol : c->opts.poisson_rtol, maxit > 0 ? maxit : c->opts.poisson_maxit,
                   iters, relres);
  HIPCHECK(c, hipMemcpy(sol, g.pot, NN * 8, hipMemcpyDeviceToHost));
  return st;
}

}  // extern "C"

// ---- the solve on a caller's workspace (emc_poisson_solve_noalloc)
// Layout, in doubles: b and x (the staged right-hand side and solution), the CG
// vectors r p p2 q z, five block-partial arrays of 3*nblocks, scal (8), the
// done flag (one slot), then per multigrid level x t r (level 0) or x t r b.
static size_t align32(size_t n) { return (n + 31) & ~(size_t)31; }  // 256-byte steps

size_t emc_poisson_ws_doubles(emc_ctx* c) {
  if (!c || !c->gpu) return 0;
  const Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  size_t need = 7 * align32(NN) + 5 * align32(3 * (size_t)g.pcg_blocks) + align32(8) + align32(1);
  if (c->opts.poisson_fp32 == 2) {
    // the FP32 cycle: r32 and z32, then x t r b per level (level 0 reads r32), in FP32
    need += 2 * align32((NN + 1) / 2);
    for (int l = 0; l < g.mg.nl; ++l) need += (l == 0 ? 3 : 4) * align32((g.mg.L[l].n + 1) / 2);
  } else {
    for (int l = 0; l < g.mg.nl; ++l) need += (l == 0 ? 3 : 4) * align32(g.mg.L[l].n);
  }
  return need;
}

int emc_ctx_device(emc_ctx* c) { return (c && c->gpu) ? c->gpu->device : -1; }

int emc_poisson_solve_ws(emc_ctx* c, const double* rhs, double* sol, double rtol, int maxit, int* iters,
                         double* relres, double* ws, size_t ws_doubles) {
  if (!c || !rhs || !sol || !ws) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t need = emc_poisson_ws_doubles(c);
  if (need > ws_doubles)
    return fail(c, EMC_ERR_CAPACITY, "Poisson workspace: the mesh needs " + std::to_string(need * 8) +
                                         " bytes, the workspace holds " + std::to_string(ws_doubles * 8));
  const size_t NN = c->dev.nnode;
  size_t off = 0;
  auto take = [&](size_t n) { double* p = ws + off; off += align32(n); return p; };
  rtol = rtol > 0 ? rtol : c->opts.poisson_rtol;
  maxit = maxit > 0 ? maxit : c->opts.poisson_maxit;
  double* b = take(NN);
  double* x = take(NN);
  PcgArgs a = pcg_args(c, b, x, rtol);
  a.r = take(NN); a.p = take(NN); a.p2 = take(NN); a.q = take(NN); a.z = take(NN);
  for (double** pp : {&a.part_a, &a.part_pq, &a.part_rr, &a.part_rz0, &a.part_rz1}) *pp = take(3 * (size_t)g.pcg_blocks);
  a.scal = take(8);
  a.done = (int*)take(1);
  HIPCHECK(c, hipMemcpyAsync(b, rhs, NN * 8, hipMemcpyHostToDevice, g.stream));
  HIPCHECK(c, hipMemcpyAsync(x, sol, NN * 8, hipMemcpyHostToDevice, g.stream));
  int st;
  if (c->opts.poisson_fp32 == 2) {
    if (!g.f32_ready) return fail(c, EMC_ERR_STATE, "the FP32 cycle is not set up");
    auto takef = [&](size_t n) { return (float*)take((n + 1) / 2); };
    a.r32 = takef(NN);
    a.z32 = takef(NN);
    MgArgsT<float> mg = g.pmg;
    mg.done = a.done;
    for (int l = 0; l < mg.nl; ++l) {
      MgLevelDevT<float>& D = mg.L[l];
      D.x = takef(D.n); D.t = takef(D.n); D.r = takef(D.n);
      D.b = l > 0 ? takef(D.n) : a.r32;
    }
    st = run_pcg_with(c, a, mg, rtol, maxit, iters, relres, false);
  } else {
    MgArgs mg = g.mg;
    mg.done = a.done;
    for (int l = 0; l < mg.nl; ++l) {
      MgLevelDev& D = mg.L[l];
      D.x = take(D.n); D.t = take(D.n); D.r = take(D.n);
      D.b = l > 0 ? take(D.n) : a.r;
    }
    st = run_pcg_with(c, a, mg, rtol, maxit, iters, relres, false);
  }
  HIPCHECK(c, hipMemcpyAsync(sol, x, NN * 8, hipMemcpyDeviceToHost, g.stream));
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  return st;
}

extern "C" {

int emc_poisson_solve_mixed(emc_ctx* c, const double* rhs, double* sol, double rtol, int maxit_outer,
                            double inner_rtol, int* outer, int* inner, double* relres) {
  if (!c || !rhs || !sol) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  HIPCHECK(c, hipMemcpy(g.rhs, rhs, NN * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.pot, sol, NN * 8, hipMemcpyHostToDevice));
  int st = run_pcg_mixed(c, g.rhs, g.pot, rtol > 0 ? rtol : c->opts.poisson_rtol, maxit_outer > 0 ? maxit_outer : 20,
                         inner_rtol > 0 ? inner_rtol : c->opts.poisson_inner_rtol, outer, inner, relres);
  HIPCHECK(c, hipMemcpy(sol, g.pot, NN * 8, hipMemcpyDeviceToHost));
  return st;
}

int emc_poisson_mg_level(emc_ctx* c, int level, int dims[3], double* a27, double* omega) {
  if (!c || !dims) return EMC_ERR_ARG;
  const std::string e = ensure_mg(c);
  if (!e.empty()) return fail(c, EMC_ERR_DECK, e);
  if (level < 0 || level >= (int)c->mg.lv.size()) return fail(c, EMC_ERR_ARG, "no such multigrid level");
  const MgLevel& L = c->mg.lv[level];
  dims[0] = L.nx; dims[1] = L.ny; dims[2] = L.nz;
  if (a27 && level > 0) std::memcpy(a27, L.a27.data(), L.a27.size() * sizeof(double));
  if (omega) *omega = L.omega;
  return EMC_OK;
}

// Fields of a step: holes from pot_h (pot2field_for_hole), electrons and the
// surface rates from pot_e (pot2field, GetSurfRoughnessPhononScRate;
// src/mcmodel.cpp:2239-2263).  Without the quantum correction both are one field.
static int launch_fields(emc_ctx* c, const double* pot_h, const double* pot_e, bool rho_from_cells) {
  Gpu& g = *c->gpu;
  if (g.hx) {
    FieldArgs h = field_args(c, pot_h, false);
    h.ex = g.hx; h.ey = g.hy; h.ez = g.hz;
    h.rates = nullptr;
    HIPCHECK(c, emc_launch_field(h, g.stream));
  }
  HIPCHECK(c, emc_launch_field(field_args(c, pot_e, rho_from_cells), g.stream));
  return EMC_OK;
}

// quantumPotential (src/mcmodel.cpp:2378-2405): out = pot, then the effective
// potential at the nodes of the quantum region when active.
static int launch_qc(emc_ctx* c, const double* pot, double* out, bool active) {
  Gpu& g = *c->gpu;
  HIPCHECK(c, hipMemcpyAsync(out, pot, c->dev.nnode * 8, hipMemcpyDeviceToDevice, g.stream));
  if (active) {
    QcArgs a = g.qa;
    a.pot = pot;
    a.qc_pot = out;
    HIPCHECK(c, emc_launch_qc_potential(a, g.stream));
  }
  return EMC_OK;
}

int emc_qc_potential(emc_ctx* c, const double* pot, double* qc_pot) {
  if (!c || !pot || !qc_pot) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  if (!g.qc_pot) return fail(c, EMC_ERR_STATE, "the deck does not set Flag_QuantumCorrection");
  const size_t NN = c->dev.nnode;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  HIPCHECK(c, hipMemcpy(g.q, pot, NN * 8, hipMemcpyHostToDevice));  // CG work vectors, free between steps
  int st = launch_qc(c, g.q, g.z, true);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  HIPCHECK(c, hipMemcpy(qc_pot, g.z, NN * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

int emc_compute_field(emc_ctx* c, const double* pot, const double* cell_rho) {
  if (!c || !pot || !cell_rho) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached");
  Gpu& g = *c->gpu;
  HIPCHECK(c, hipMemcpy(g.pot, pot, c->dev.nnode * 8, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(g.rho, cell_rho, c->dev.ncell * 8, hipMemcpyHostToDevice));
  int st = launch_fields(c, g.pot, g.pot, false);
  if (st) return st;
  HIPCHECK(c, hipStreamSynchronize(g.stream));
  return EMC_OK;
}

int emc_potential_get(emc_ctx* c, double* pot) {
  if (!c || !pot || !c->gpu) return EMC_ERR_ARG;
  HIPCHECK(c, hipStreamSynchronize(c->gpu->stream));
  HIPCHECK(c, hipMemcpy(pot, c->gpu->pot, c->dev.nnode * 8, hipMemcpyDeviceToHost));
  return EMC_OK;
}

// statistic()'s window sums of the step; in the heat phase the node heat is
// cleared once added (heat_density, src/heat.cpp:48-66).
static StatArgs stat_args(emc_ctx* c, bool heat_phase) {
  Gpu& g = *c->gpu;
  StatArgs a{};
  a.nnode = c->dev.nnode; a.ncell = c->dev.ncell;
  a.pot = g.pot; a.stat_pot = g.stat_pot;
  a.cell_n = g.cell_n; a.stat_enum = g.stat_enum;
  if (g.heat_live || heat_phase) {
    a.heat_e = g.heat_e; a.heat_h = g.heat_h;
    a.stat_heat_e = g.stat_heat_e; a.stat_heat_h = g.stat_heat_h;
  }
  a.qc_pot = g.qc_pot; a.stat_qc_pot = g.stat_qc_pot;
  a.zero_heat = heat_phase ? 1 : 0;
  return a;
}

// One step of density() + update_particle (src/mcmodel.cpp:424-471), or with
// heat_phase one of heat_density (src/heat.cpp:48-98): deposit and window sums,
// then the Poisson solve (transient only), fields and surface rates, the flight
// and its clone passes, MultipleRefresh every mr_step steps.
static int one_step(emc_ctx* c, bool heat_phase, emc_fly_tallies* acc, double tsum[5]) {
  Gpu& g = *c->gpu;
  const size_t NN = c->dev.nnode;
  const Config& cf = c->dev.cfg;
  {
    int st = settle_n(c);
    if (st) return st;
    st = grow_if_full(c);
    if (st) return st;
  }
  const uint64_t count = heat_phase ? c->heat_next++ : c->next_step++;
  const uint64_t step = heat_phase ? HEAT_STEP0 + count : count;
  HIPCHECK(c, hipEventRecord(g.ev[0], g.stream));
  int st = run_sort(c);
  if (st) return st;
  // the deposit's node gather also forms the right-hand side (transient) and the
  // window sums
  DepositArgs da = deposit_args(c, 1);
  da.st = stat_args(c, heat_phase);
  if (!heat_phase) da.rhs = RhsArgs{NN, g.node_q, g.pvol, g.pdop, g.dirichlet, g.corr, g.kind, g.rhs, nullptr};
  HIPCHECK(c, emc_launch_deposit(da, g.stream));
  HIPCHECK(c, hipEventRecord(g.ev[1], g.stream));
  if (!heat_phase) {
    if (c->opts.poisson_fp32 == 1)
      st = run_pcg_mixed(c, g.rhs, g.pot, c->opts.poisson_rtol, 20, c->opts.poisson_inner_rtol, nullptr, nullptr,
                         nullptr);
    else
      st = run_pcg(c, g.rhs, g.pot, c->opts.poisson_rtol, c->opts.poisson_maxit, nullptr, nullptr, true);
    if (st) return st;
  }
  HIPCHECK(c, hipEventRecord(g.ev[2], g.stream));
  if (heat_phase) {
    // heat_field (src/heat.cpp:100-103): holes at the window-mean Poisson
    // potential, electrons at the window-mean corrected one
    st = launch_fields(c, g.pot_avg, g.qc_avg ? g.qc_avg : g.pot_avg, true);
  } else {
    const double* pe = g.pot;
    if (g.qc_pot) {
      st = launch_qc(c, g.pot, g.qc_pot, count >= (uint64_t)std::max(0, cf.quantum_start_step));
      if (st) return st;
      pe = g.qc_pot;
    }
    st = launch_fields(c, g.pot, pe, true);
  }
  if (st) return st;
  HIPCHECK(c, hipEventRecord(g.ev[3], g.stream));
  st = zero_tallies(c);
  if (st) return st;
  FlyArgs fa{};
  // the transient flies without heat tallies: run() clears flag_heat (src/mcmodel.cpp:515)
  st = launch_fly(c, step, heat_phase ? 1 : 0, fa);
  if (st) return st;
  HIPCHECK(c, hipEventRecord(g.ev[4], g.stream));
  HIPCHECK(c, hipEventSynchronize(g.ev[4]));
  // the one wait of the step: the solve's result, then the flight's counts
  if (!heat_phase) {
    st = finish_pcg(c, c->opts.poisson_rtol, nullptr, nullptr);
    if (st) return st;
  }
  const int passes = finish_fly(c, fa);
  if (passes < 0) return passes;
  // MultipleRefresh after the flight of every mr_step-th step (src/mcmodel.cpp:535-547)
  long long mr_gen = 0;
  if (cf.flag_mr && cf.mr_step > 0 && count % (uint64_t)cf.mr_step == (uint64_t)cf.mr_step - 1) {
    long long nb = 0, na = 0;
    st = run_mr(c, step, &nb, &na);
    if (st) return st;
    mr_gen = na - nb;
  }
  float ms;
  for (int q = 0; q < 4; ++q) {
    HIPCHECK(c, hipEventElapsedTime(&ms, g.ev[q], g.ev[q + 1]));
    tsum[q] += ms;
  }
  HIPCHECK(c, hipEventElapsedTime(&ms, g.ev[5], g.ev[6]));  // the k_fly launch alone
  tsum[4] += ms;
  read_tallies(c, acc, passes);
  read_tallies(c, &c->win, passes);
  if (acc) acc->n_mr_gen += mr_gen;
  c->win.n_mr_gen += mr_gen;
  ++c->win_steps;
  return EMC_OK;
}

static int run_steps(emc_ctx* c, int nsteps, emc_fly_tallies* acc, bool heat_phase) {
  double tsum[5] = {0, 0, 0, 0, 0};
  for (int s = 0; s < nsteps; ++s) {
    const int st = one_step(c, heat_phase, acc, tsum);
    if (st) return st;
  }
  for (int q = 0; q < 5; ++q) c->timing[q] = nsteps ? tsum[q] / nsteps : 0;
  return EMC_OK;
}

// density() + update_particle for nsteps (src/mcmodel.cpp:424-471): deposit the
// current population, solve Poisson (warm start), fields and surface rates,
// then fly.  Window statistics accumulate across calls until emc_tallies_reset.
int emc_step(emc_ctx* c, int nsteps, emc_fly_tallies* acc) {
  if (!c || nsteps < 0) return EMC_ERR_ARG;
  if (!c->gpu) return fail(c, EMC_ERR_NOGPU, "no GPU attached: the step runs on the GPU only");
  if (c->gpu->cap == 0) return fail(c, EMC_ERR_STATE, "no particles on the device");
  return run_steps(c, nsteps, acc, false);
}

int emc_step_timing(emc_ctx* c, double* ms5, int* pcg_iters) {
  if (!c) return EMC_ERR_ARG;
  if (ms5)
    for (int q = 0; q < 5; ++q) ms5[q] = c->timing[q];
  if (pcg_iters) *pcg_iters = c->last_pcg_iters;
  return EMC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// A batch of K devices that share the mesh (the same deck geometry; contact
// voltages, particle counts and seeds may differ), stepped together: each kernel
// of the step is one launch for all of them (gridDim.y = K, the argument blocks
// in device tables), so a bias sweep of small devices fills the GPU with one
// stream of launches instead of K.  Each device's blocks run the code and grids
// of its own launch, so every device's result equals its own emc_step bit for
// bit.  The host waits on the GPU as one device does: the solve's residual checks
// and once at the end of the step, for all K at once.
struct emc_batch {
  std::vector<emc_ctx*> c;
  int K = 0, device = -1;
  hipStream_t s = nullptr;
  std::string err;
  DevBuf mem;
  // device argument tables: the per-step part (sort, deposit) and the flight
  // are uploaded each step, the rest once
  unsigned char* d_tab = nullptr;
  unsigned char* h_tab = nullptr;     // pinned host image of d_tab
  size_t off_sort = 0, off_dep = 0, off_pcg = 0, off_mg = 0, off_fmg = 0, off_field = 0, off_fly = 0, off_cp = 0,
         tab_bytes = 0;
  BatchTables bt{};
  const CopyArgs *cp_n = nullptr, *cp_scal = nullptr, *cp_tally = nullptr, *cp_zero = nullptr;
  unsigned long long* d_n = nullptr;   // [K] live counts after the sort
  double* d_scal = nullptr;            // [2][K][8] PCG scalars of the last two iterations
  unsigned long long* d_tally = nullptr;  // [K][FLYOUT_N]
  unsigned long long* h_n = nullptr;
  double* h_scal = nullptr;
  unsigned long long* h_tally = nullptr;
  hipEvent_t ev[7] = {};
  hipEvent_t ev_scal[2] = {nullptr, nullptr};
  hipEvent_t ev_n = nullptr;
  MgArgs shape{};
  int nblocks = 0, ntile = 0;
  size_t nnode = 0, ncell = 0;
  double timing[5] = {0, 0, 0, 0, 0};
  int last_iters = 0;
};

static int bfail(emc_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}
#define BHIP(b, expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return bfail(b, EMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

static size_t tab_align(size_t n) { return (n + 255) & ~(size_t)255; }

static int batch_setup(emc_batch* b) {
  const int K = b->K;
  emc_ctx* c0 = b->c[0];
  for (emc_ctx* c : b->c) {
    if (!c || !c->gpu) return bfail(b, EMC_ERR_NOGPU, "every context of a batch needs a GPU");
    if (c->gpu->cap == 0) return bfail(b, EMC_ERR_STATE, "a context of the batch has no particles on the device");
    if (c->gpu->device != c0->gpu->device) return bfail(b, EMC_ERR_ARG, "the contexts of a batch share one GPU");
    const Device &d = c->dev, &d0 = c0->dev;
    if (d.cnx != d0.cnx || d.cny != d0.cny || d.cnz != d0.cnz)
      return bfail(b, EMC_ERR_ARG, "the contexts of a batch share the mesh dimensions");
    if (c->gpu->mg.nl != c0->gpu->mg.nl || c->gpu->mg.nc != c0->gpu->mg.nc)
      return bfail(b, EMC_ERR_ARG, "the contexts of a batch need the same multigrid hierarchy shape");
    for (int l = 0; l < c0->gpu->mg.nl; ++l)
      if (c->gpu->mg.L[l].n != c0->gpu->mg.L[l].n)
        return bfail(b, EMC_ERR_ARG, "the contexts of a batch need the same multigrid hierarchy shape");
    if (d.cfg.flag_qc) return bfail(b, EMC_ERR_ARG, "batches run without the quantum correction");
    if (c->opts.poisson_rtol != c0->opts.poisson_rtol || c->opts.poisson_maxit != c0->opts.poisson_maxit)
      return bfail(b, EMC_ERR_ARG, "the contexts of a batch share the solver tolerance");
    if (c->opts.poisson_fp32 == 1 || c->opts.poisson_fp32 != c0->opts.poisson_fp32)
      return bfail(b, EMC_ERR_ARG, "the contexts of a batch share one FP64 CG (poisson_fp32 0 or 2 for all)");
    if (c->heat_phase) return bfail(b, EMC_ERR_STATE, "a context of the batch is in its heating phase");
    for (emc_ctx* o : b->c)
      if (o != c && o->gpu.get() == c->gpu.get()) return bfail(b, EMC_ERR_ARG, "a context appears twice");
  }
  for (int q = 0; q < K; ++q)
    for (int r = 0; r < q; ++r)
      if (b->c[q] == b->c[r]) return bfail(b, EMC_ERR_ARG, "a context appears twice in the batch");
  b->device = c0->gpu->device;
  BHIP(b, hipSetDevice(b->device));
  for (emc_ctx* c : b->c) BHIP(b, hipStreamSynchronize(c->gpu->stream));
  BHIP(b, hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
  for (auto& e : b->ev) BHIP(b, hipEventCreate(&e));
  for (auto& e : b->ev_scal) BHIP(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  BHIP(b, hipEventCreateWithFlags(&b->ev_n, hipEventDisableTiming));
  const Gpu& g0 = *c0->gpu;
  b->shape = g0.mg;
  b->nblocks = g0.pcg_blocks;
  b->ntile = g0.ntile;
  b->nnode = c0->dev.nnode;
  b->ncell = c0->dev.ncell;
  size_t off = 0;
  auto place = [&](size_t bytes) { const size_t o = off; off += tab_align(bytes); return o; };
  b->off_sort = place(K * sizeof(SortArgs));
  b->off_dep = place(K * sizeof(DepositArgs));
  b->off_fly = place(K * sizeof(FlyArgs));
  b->off_pcg = place(K * sizeof(PcgArgs));
  b->off_mg = place(3 * K * sizeof(MgArgs));
  const bool f32 = c0->opts.poisson_fp32 == 2;
  b->off_fmg = place(f32 ? 3 * K * sizeof(MgArgsT<float>) : 0);
  b->off_field = place(K * sizeof(FieldArgs));
  b->off_cp = place(5 * K * sizeof(CopyArgs));
  b->tab_bytes = off;
  b->d_tab = b->mem.alloc<unsigned char>(off);
  b->d_n = b->mem.alloc<unsigned long long>(K);
  b->d_scal = b->mem.alloc<double>(2 * 8 * (size_t)K);
  b->d_tally = b->mem.alloc<unsigned long long>((size_t)K * FLYOUT_N);
  if (b->mem.last != hipSuccess) return bfail(b, EMC_ERR_HIP, std::string("batch tables: ") + hipGetErrorString(b->mem.last));
  BHIP(b, hipHostMalloc((void**)&b->h_tab, off, hipHostMallocDefault));
  BHIP(b, hipHostMalloc((void**)&b->h_n, K * sizeof(unsigned long long), hipHostMallocDefault));
  BHIP(b, hipHostMalloc((void**)&b->h_scal, 2 * 8 * (size_t)K * sizeof(double), hipHostMallocDefault));
  BHIP(b, hipHostMalloc((void**)&b->h_tally, (size_t)K * FLYOUT_N * sizeof(unsigned long long), hipHostMallocDefault));
  std::memset(b->h_tab, 0, off);
  b->bt.K = K;
  b->bt.sort = (const SortArgs*)(b->d_tab + b->off_sort);
  b->bt.dep = (const DepositArgs*)(b->d_tab + b->off_dep);
  b->bt.fly = (const FlyArgs*)(b->d_tab + b->off_fly);
  b->bt.pcg = (const PcgArgs*)(b->d_tab + b->off_pcg);
  const MgArgs* mg = (const MgArgs*)(b->d_tab + b->off_mg);
  b->bt.mg_start = mg;
  b->bt.mg_even = mg + K;
  b->bt.mg_odd = mg + 2 * K;
  if (f32) {
    const MgArgsT<float>* fm = (const MgArgsT<float>*)(b->d_tab + b->off_fmg);
    b->bt.fmg_start = fm;
    b->bt.fmg_even = fm + K;
    b->bt.fmg_odd = fm + 2 * K;
  }
  b->bt.field = (const FieldArgs*)(b->d_tab + b->off_field);
  const CopyArgs* cp = (const CopyArgs*)(b->d_tab + b->off_cp);
  b->cp_n = cp;
  b->cp_scal = cp + K;  // [2][K]
  b->cp_tally = cp + 3 * K;
  b->cp_zero = cp + 4 * K;
  // the tables that stay: solver, cycle, fields, readbacks
  PcgArgs* hp = (PcgArgs*)(b->h_tab + b->off_pcg);
  MgArgs* hm = (MgArgs*)(b->h_tab + b->off_mg);
  MgArgsT<float>* hfm = (MgArgsT<float>*)(b->h_tab + b->off_fmg);
  FieldArgs* hf = (FieldArgs*)(b->h_tab + b->off_field);
  CopyArgs* hc = (CopyArgs*)(b->h_tab + b->off_cp);
  for (int q = 0; q < K; ++q) {
    emc_ctx* c = b->c[q];
    Gpu& g = *c->gpu;
    hp[q] = pcg_args(c, g.rhs, g.pot, c->opts.poisson_rtol);
    hm[q] = emc_mg_start_args(hp[q], g.mg);
    hm[K + q] = emc_mg_iter_args(hp[q], g.mg, 0);
    hm[2 * K + q] = emc_mg_iter_args(hp[q], g.mg, 1);
    if (f32) {
      if (!g.f32_ready) return bfail(b, EMC_ERR_STATE, "a context of the batch has no FP32 cycle");
      hp[q].r32 = g.r32;
      hp[q].z32 = g.z32;
      hfm[q] = emc_mg_start_args(hp[q], g.pmg);
      hfm[K + q] = emc_mg_iter_args(hp[q], g.pmg, 0);
      hfm[2 * K + q] = emc_mg_iter_args(hp[q], g.pmg, 1);
    }
    hf[q] = field_args(c, g.pot, true);
    int st = sync_fly_const(c);
    if (st) return bfail(b, st, c->err);
    hc[q] = CopyArgs{g.toff + g.ntile, b->d_n + q, 1, 1, 0};  // the live count, a 32-bit word
    hc[K + q] = CopyArgs{g.scal, (unsigned long long*)(b->d_scal + 8 * (size_t)q), 8, 0, 0};
    hc[2 * K + q] = CopyArgs{g.scal, (unsigned long long*)(b->d_scal + 8 * ((size_t)K + q)), 8, 0, 0};
    hc[3 * K + q] = CopyArgs{g.contact, b->d_tally + (size_t)q * FLYOUT_N, FLYOUT_N, 0, 0};
    hc[4 * K + q] = CopyArgs{nullptr, g.contact, FLYOUT_N, 0, 0};
  }
  BHIP(b, hipMemcpyAsync(b->d_tab + b->off_pcg, b->h_tab + b->off_pcg, b->tab_bytes - b->off_pcg,
                         hipMemcpyHostToDevice, b->s));
  BHIP(b, hipStreamSynchronize(b->s));
  return EMC_OK;
}

// One step of every device of the batch: one_step's phases, each one launch for all.
static int batch_one_step(emc_batch* b, emc_fly_tallies* acc, double tsum[5], int* iters_max) {
  const int K = b->K;
  hipStream_t s = b->s;
  std::vector<uint64_t> step(K);
  SortArgs* hs = (SortArgs*)(b->h_tab + b->off_sort);
  DepositArgs* hd = (DepositArgs*)(b->h_tab + b->off_dep);
  FlyArgs* hfl = (FlyArgs*)(b->h_tab + b->off_fly);
  long long nmax = 0;
  int guess = 0;
  for (int q = 0; q < K; ++q) {
    emc_ctx* c = b->c[q];
    Gpu& g = *c->gpu;
    step[q] = c->next_step++;
    int st = settle_n(c);
    if (st) return bfail(b, st, c->err);
    st = grow_if_full(c);
    if (st) return bfail(b, st, c->err);
    hs[q] = sort_args(c);
    nmax = std::max(nmax, sort_nappend(c));
    g.seg_valid = true;
    std::swap(g.P, g.P2);
    hd[q] = deposit_args(c, 1);
    hd[q].st = stat_args(c, false);
    hd[q].rhs = RhsArgs{b->nnode, g.node_q, g.pvol, g.pdop, g.dirichlet, g.corr, g.kind, g.rhs, nullptr};
    guess = std::max(guess, c->last_pcg_iters);
  }
  BHIP(b, hipEventRecord(b->ev[0], s));
  BHIP(b, hipMemcpyAsync(b->d_tab + b->off_sort, b->h_tab + b->off_sort, b->off_fly - b->off_sort,
                         hipMemcpyHostToDevice, s));
  BHIP(b, emc_launch_tile_sort_b(b->bt, b->ntile, nmax, s));
  // the live counts (32-bit toff[ntile] of each device) into d_n
  BHIP(b, emc_launch_copy_b(b->cp_n, K, 1, s));
  BHIP(b, hipMemcpyAsync(b->h_n, b->d_n, K * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  BHIP(b, hipEventRecord(b->ev_n, s));
  BHIP(b, emc_launch_deposit_b(b->bt, b->ntile, b->nnode, s));
  BHIP(b, hipEventRecord(b->ev[1], s));
  // the solve: as run_pcg, one iteration of every device per launch; the host
  // checks the iteration before the one just queued, for all devices at once
  BHIP(b, emc_launch_pcg_start_b(b->bt, b->shape, b->nblocks, s));
  const int maxit = b->c[0]->opts.poisson_maxit;
  int checked = -1, last = -1;
  for (int k = 0; k < maxit; ++k) {
    last = k;
    BHIP(b, emc_launch_pcg_iter_b(b->bt, b->shape, b->nblocks, k, s));
    BHIP(b, emc_launch_copy_b(b->cp_scal + (size_t)(k & 1) * K, K, 8, s));
    BHIP(b, hipMemcpyAsync(b->h_scal + (size_t)(k & 1) * 8 * K, b->d_scal + (size_t)(k & 1) * 8 * K,
                           8 * K * sizeof(double), hipMemcpyDeviceToHost, s));
    BHIP(b, hipEventRecord(b->ev_scal[k & 1], s));
    const int chk = (k + 1 == guess) ? k : k - 1;
    if (chk < 0 || chk <= checked) continue;
    BHIP(b, hipEventSynchronize(b->ev_scal[chk & 1]));
    checked = chk;
    bool all = true;
    for (int q = 0; q < K; ++q) {
      const double* res = b->h_scal + (size_t)(chk & 1) * 8 * K + 8 * q;
      if (!std::isfinite(res[0])) return bfail(b, EMC_ERR_SOLVER, "PCG residual is not finite");
      if (!(res[0] <= res[4])) all = false;
    }
    if (all) break;
  }
  BHIP(b, hipEventRecord(b->ev[2], s));
  BHIP(b, emc_launch_field_b(b->bt, b->ncell, s));
  BHIP(b, hipEventRecord(b->ev[3], s));
  BHIP(b, emc_launch_copy_b(b->cp_zero, K, FLYOUT_N, s));
  // the flight needs the live counts of the sort (long since read: the residual
  // checks waited on later work)
  BHIP(b, hipEventSynchronize(b->ev_n));
  nmax = 0;
  for (int q = 0; q < K; ++q) {
    emc_ctx* c = b->c[q];
    Gpu& g = *c->gpu;
    g.n = (long long)(unsigned)b->h_n[q];
    g.sorted_n = g.n;
    g.n_pending = false;
    hfl[q] = fly_args(c, step[q], 0);
    nmax = std::max(nmax, g.n);
  }
  BHIP(b, hipMemcpyAsync(b->d_tab + b->off_fly, b->h_tab + b->off_fly, K * sizeof(FlyArgs), hipMemcpyHostToDevice, s));
  BHIP(b, hipEventRecord(b->ev[5], s));
  BHIP(b, emc_launch_fly_b(b->bt, nmax, b->c[0]->gpu->wave_slots, s));
  BHIP(b, hipEventRecord(b->ev[6], s));
  BHIP(b, emc_launch_copy_b(b->cp_tally, K, FLYOUT_N, s));
  BHIP(b, hipMemcpyAsync(b->h_tally, b->d_tally, (size_t)K * FLYOUT_N * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, s));
  BHIP(b, hipEventRecord(b->ev[4], s));
  BHIP(b, hipEventSynchronize(b->ev[4]));
  // per device: the solve's result, the flight's counts (finish_fly), the window
  int imax = 0;
  for (int q = 0; q < K; ++q) {
    emc_ctx* c = b->c[q];
    Gpu& g = *c->gpu;
    if (last >= 0) {
      const double* sh = b->h_scal + (size_t)(last & 1) * 8 * K + 8 * q;
      const double bn = std::sqrt(sh[2]) > 0 ? std::sqrt(sh[2]) : 1.0;
      const double rel = std::sqrt(std::max(0.0, sh[0])) / bn;
      c->last_pcg_iters = (int)sh[3];
      c->last_relres = rel;
      imax = std::max(imax, (int)sh[3]);
      if (!std::isfinite(rel)) return bfail(b, EMC_ERR_SOLVER, "PCG residual is not finite");
      if (rel > c->opts.poisson_rtol) return bfail(b, EMC_ERR_SOLVER, "PCG did not reach the tolerance");
    }
    std::memcpy(g.tally_host, b->h_tally + (size_t)q * FLYOUT_N, FLYOUT_N * sizeof(unsigned long long));
    const int passes = finish_fly(c, hfl[q]);
    if (passes < 0) return bfail(b, passes, c->err);
    const Config& cf = c->dev.cfg;
    long long mr_gen = 0;
    if (cf.flag_mr && cf.mr_step > 0 && step[q] % (uint64_t)cf.mr_step == (uint64_t)cf.mr_step - 1) {
      long long nb = 0, na = 0;
      const int st = run_mr(c, step[q], &nb, &na);
      if (st) return bfail(b, st, c->err);
      mr_gen = na - nb;
    }
    read_tallies(c, acc ? &acc[q] : nullptr, passes);
    read_tallies(c, &c->win, passes);
    if (acc) acc[q].n_mr_gen += mr_gen;
    c->win.n_mr_gen += mr_gen;
    ++c->win_steps;
    // work finish_fly or the refresh queued on the device's own stream
    BHIP(b, hipStreamSynchronize(g.stream));
  }
  *iters_max = imax;
  float ms;
  for (int q = 0; q < 3; ++q) {
    BHIP(b, hipEventElapsedTime(&ms, b->ev[q], b->ev[q + 1]));
    tsum[q] += ms;
  }
  BHIP(b, hipEventElapsedTime(&ms, b->ev[3], b->ev[4]));
  tsum[3] += ms;
  BHIP(b, hipEventElapsedTime(&ms, b->ev[5], b->ev[6]));
  tsum[4] += ms;
  return EMC_OK;
}

extern "C" {

emc_batch* emc_batch_create(emc_ctx* const* ctx, int n, int* status) {
  int st = EMC_OK;
  emc_batch* b = nullptr;
  if (!ctx || n < 1) st = EMC_ERR_ARG;
  else {
    b = new emc_batch();
    b->K = n;
    b->c.assign(ctx, ctx + n);
    st = batch_setup(b);
  }
  if (status) *status = st;
  return b;
}

const char* emc_batch_last_error(emc_batch* b) { return b ? b->err.c_str() : "null batch"; }

int emc_batch_step(emc_batch* b, int nsteps, emc_fly_tallies* acc) {
  if (!b || nsteps < 0) return EMC_ERR_ARG;
  if (!b->s) return bfail(b, EMC_ERR_STATE, "the batch was not set up: " + b->err);
  for (emc_ctx* c : b->c)
    if (c->heat_phase) return bfail(b, EMC_ERR_STATE, "a context of the batch is in its heating phase");
  double tsum[5] = {0, 0, 0, 0, 0};
  int imax = 0;
  for (int q = 0; q < nsteps; ++q) {
    const int st = batch_one_step(b, acc, tsum, &imax);
    if (st) return st;
  }
  for (int q = 0; q < 5; ++q) b->timing[q] = nsteps ? tsum[q] / nsteps : 0;
  b->last_iters = imax;
  return EMC_OK;
}

int emc_batch_timing(emc_batch* b, double* ms5, int* pcg_iters_max) {
  if (!b) return EMC_ERR_ARG;
  if (ms5)
    for (int q = 0; q < 5; ++q) ms5[q] = b->timing[q];
  if (pcg_iters_max) *pcg_iters_max = b->last_iters;
  return EMC_OK;
}

void emc_batch_destroy(emc_batch* b) {
  if (!b) return;
  if (b->device >= 0) {
    (void)hipSetDevice(b->device);
    if (b->s) (void)hipStreamSynchronize(b->s);
  }
  for (auto e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto e : b->ev_scal)
    if (e) (void)hipEventDestroy(e);
  if (b->ev_n) (void)hipEventDestroy(b->ev_n);
  if (b->h_tab) (void)hipHostFree(b->h_tab);
  if (b->h_n) (void)hipHostFree(b->h_n);
  if (b->h_scal) (void)hipHostFree(b->h_scal);
  if (b->h_tally) (void)hipHostFree(b->h_tally);
  if (b->s) (void)hipStreamDestroy(b->s);
  b->mem.free_all();
  delete b;
}

int emc_run_file(const char* input_txt, const emc_opts* o) {
  int st = 0;
  emc_ctx* c = emc_create(input_txt, o, &st);
  auto stop = [&](const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, emc_last_error(c));
    emc_destroy(c);
    return st;
  };
  if (st) return stop("emc_create");
  if (!c->gpu) {
    st = fail(c, EMC_ERR_NOGPU, "a run needs a GPU (opts.device >= 0)");
    return stop("emc_create");
  }
  int64_t n = 0;
  if ((st = emc_particles_init(c, &n))) return stop("initial population");
  const Config& cf = c->dev.cfg;
  if (cf.stat_step <= 0 || (cf.flag_heat && cf.stat_heat_step <= 0)) {
    st = fail(c, EMC_ERR_DECK, "stat_step and stat_heat_step must be positive");
    return stop("deck");
  }
  const char* dir = "data";
  const int s0 = cf.flag_restart ? std::max(0, cf.restart_step) : 0;
  std::printf("Entering run(): total_step = %d, restart step = %d\n", cf.total_step, cf.restart_step);
  if ((st = emc_output_pvolume(c, dir))) return stop("data/pvolume");
  // run() (src/mcmodel.cpp:498-566)
  for (int s = s0; s < cf.total_step; ++s) {
    std::printf("************ step = %d ********\n%lld particles\n", s, (long long)c->gpu->n);
    emc_fly_tallies t;
    std::memset(&t, 0, sizeof(t));
    if ((st = emc_step(c, 1, &t))) return stop("step");
    if (cf.flag_mr && cf.mr_step > 0 && s % cf.mr_step == cf.mr_step - 1)
      std::printf("MultipleRefresh Gen : %lld\n", (long long)t.n_mr_gen);
    if (s % cf.stat_step == cf.stat_step - 1) {
      if ((st = emc_output_stat(c, dir, s)) || (st = emc_output_for_restart(c, dir, s)) || (st = emc_tallies_reset(c)))
        return stop("window output");
    }
  }
  // compute_heat (src/heat.cpp:108-167)
  if (cf.flag_heat) {
    std::printf("begin computing heating for fixed potential\n");
    if ((st = emc_heat_begin(c))) return stop("heat phase");
    for (int s = 0; s < cf.heat_steps; ++s) {
      std::printf("************ step = %d ********\n%lld particles\n", s, (long long)c->gpu->n);
      if ((st = emc_heat_step(c, 1, nullptr))) return stop("heat step");
      if (s % cf.stat_heat_step == cf.stat_heat_step - 1) {
        if ((st = emc_output_stat(c, dir, s)) || (st = emc_output_for_restart(c, dir, s)) ||
            (st = emc_tallies_reset(c)))
          return stop("window output");
      }
    }
  }
  std::fflush(stdout);
  emc_destroy(c);
  return EMC_OK;
}

}  // extern "C"
