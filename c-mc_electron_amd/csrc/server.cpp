// emc_server: a serving loop over the C ABI.  Requests (a deck and a target
// precision for the drain current) arrive at any time; a worker thread admits
// them by their estimated device-memory footprint against a budget, keeps their
// state resident on the GPU, and steps every admitted request: requests with the
// same mesh dimensions step together as one batch (emc_batch: one launch per
// kernel for all of them), different meshes in batches of their own on their own
// streams at the same time.  Admission, grouping and completion happen between
// chunks of steps that end on every request's window boundary, so a request that
// arrives while others run joins at the next chunk.
//
// A request runs windows of window_steps steps; each window gives one drain
// current, the mean over contacts 0 and 1 of |Current| as output_stat writes it
// (src/mcmodel.cpp:756-842; finfet/scripts/current.py averages these).  After
// skip_windows windows the request is done when the Student-t interval of the
// mean over the later windows is within rel_tol of the mean (at least
// min_windows of them), or after max_windows.  A request's result depends only
// on its deck, seed and window settings: a batch steps every device exactly as
// it would step alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/emc.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Student-t 0.975 quantiles for 1..30 degrees of freedom, then the normal value.
double t975(int df) {
  static const double t[31] = {0,      12.706, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365, 2.306, 2.262, 2.228,
                               2.201,  2.179,  2.160, 2.145, 2.131, 2.120, 2.110, 2.101, 2.093, 2.086, 2.080,
                               2.074,  2.069,  2.064, 2.060, 2.056, 2.052, 2.048, 2.045, 2.042};
  if (df < 1) return INFINITY;
  return df <= 30 ? t[df] : 1.960 + 2.4 / df;
}

struct Req {
  int64_t id = 0;
  std::string deck;
  emc_request_opts o{};
  emc_ctx* c = nullptr;
  int state = EMC_REQ_QUEUED;
  int status = EMC_OK;
  std::string err;
  double t_submit = 0, t_start = 0, t_done = 0;
  int dims[3] = {0, 0, 0};
  double dt = 0, curr0 = 0;
  double est_bytes = 0, bytes = 0;
  int64_t steps = 0, particles = 0, nodes = 0;
  int win_step = 0;
  emc_fly_tallies acc{};
  std::vector<double> win;  // drain current of each finished window
  double mean = 0, half = 0;
  bool converged = false;
};

// key=value of input.txt ('#' starts a comment)
double deck_number(const std::string& path, const char* key, double dflt) {
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string k = line.substr(0, eq);
    k.erase(std::remove_if(k.begin(), k.end(), ::isspace), k.end());
    if (k == key) return std::atof(line.c_str() + eq + 1);
  }
  return dflt;
}

}  // namespace

struct emc_server {
  int device = 0;
  double budget = 0;
  int max_active = 0;
  std::string table_dir;
  emc_opts base{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req*> queue;
  std::vector<Req*> active;
  std::map<int64_t, std::unique_ptr<Req>> all;
  int64_t next_id = 1;
  double used = 0;
  bool stop = false;
  std::thread worker;
  double t0 = 0;
  std::string err;
};

namespace {

// Device bytes a request will hold, from its deck: per node the tallies, window
// sums, potential, solver and multigrid vectors; per cell the fields, rates, cell
// sums and deposit tile blocks; per particle slot the two particle buffers and the
// flight's work lists; the dense inverse of the coarsest multigrid level.
double estimate_bytes(const emc_mesh_view& m, const std::string& deck, double capacity_factor) {
  const double ne = deck_number(deck, "ElectronNumber", 0), nh = deck_number(deck, "HoleNumber", 0);
  const double slots = (ne + nh + 2.0 * (double)m.ncell) * std::max(1.0, capacity_factor) + 1024;
  const double nc = std::max(600.0, std::min(2500.0, (double)m.nnode / 300));
  return 320.0 * (double)m.nnode + 360.0 * (double)m.ncell + 190.0 * slots + 8.0 * nc * nc + (8 << 20);
}

void finish_window(Req* r) {
  // output_stat's Current of contact q (src/mcmodel.cpp:811)
  auto current = [&](int q) {
    const double gen = (double)r->acc.contact[q][EMC_CT_QGEN] / EMC_FX_CONTACT;
    const double cat = (double)r->acc.contact[q][EMC_CT_QCATCH] / EMC_FX_CONTACT;
    return (gen - cat) / (r->dt * r->o.window_steps) * r->curr0 * 1e-2;
  };
  if (r->o.drain_contact >= 0)
    r->win.push_back(current(r->o.drain_contact));
  else
    r->win.push_back(0.5 * (std::fabs(current(0)) + std::fabs(current(1))));
  std::memset(&r->acc, 0, sizeof(r->acc));
  r->win_step = 0;
  const int skip = std::max(0, r->o.skip_windows);
  const int n = (int)r->win.size() - skip;
  if (n < 1) return;
  double m = 0;
  for (int q = skip; q < (int)r->win.size(); ++q) m += r->win[q];
  m /= n;
  double v = 0;
  for (int q = skip; q < (int)r->win.size(); ++q) v += (r->win[q] - m) * (r->win[q] - m);
  const double sd = n > 1 ? std::sqrt(v / (n - 1)) : INFINITY;
  r->mean = m;
  r->half = n > 1 ? t975(n - 1) * sd / std::sqrt((double)n) : INFINITY;
  r->converged = n >= std::max(2, r->o.min_windows) && r->half <= std::max(r->o.rel_tol * std::fabs(m), r->o.abs_tol);
}

void fail_req(Req* r, int st, const std::string& why) {
  r->status = st;
  r->err = why;
}

// Create the request's context on the GPU and its initial population.
void start_request(emc_server* s, Req* r) {
  emc_opts o = s->base;
  o.device = s->device;
  o.seed = r->o.seed;
  o.table_dir = s->table_dir.c_str();
  int st = 0;
  size_t f0 = 0, f1 = 0, tot = 0;
  (void)hipSetDevice(s->device);
  (void)hipMemGetInfo(&f0, &tot);
  r->c = emc_create(r->deck.c_str(), &o, &st);
  if (st) {
    fail_req(r, st, emc_last_error(r->c));
    return;
  }
  int64_t n = 0;
  if ((st = emc_particles_init(r->c, &n)) || (st = emc_tallies_reset(r->c))) {
    fail_req(r, st, emc_last_error(r->c));
    return;
  }
  (void)hipMemGetInfo(&f1, &tot);
  // reported only: the budget is kept on the estimates, which do not move when
  // memory another request released is returned late
  r->bytes = f0 > f1 ? (double)(f0 - f1) : r->est_bytes;
  r->particles = n;
}

void run_worker(emc_server* s) {
  std::map<std::vector<Req*>, emc_batch*> batches;  // by member list
  auto drop_batches = [&]() {
    for (auto& kv : batches) emc_batch_destroy(kv.second);
    batches.clear();
  };
  for (;;) {
    std::vector<Req*> admit;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->active.empty(); });
      if (s->stop) break;
      // admission by footprint, in arrival order
      while (!s->queue.empty() && (int)(s->active.size() + admit.size()) < s->max_active) {
        Req* r = s->queue.front();
        const bool alone = s->active.empty() && admit.empty();
        if (!alone && s->used + r->est_bytes > s->budget) break;
        s->queue.pop_front();
        s->used += r->est_bytes;
        admit.push_back(r);
      }
    }
    for (Req* r : admit) {
      r->t_start = now_s() - s->t0;
      start_request(s, r);
      std::lock_guard<std::mutex> lk(s->mu);
      if (r->status) {
        r->state = EMC_REQ_FAILED;
        r->t_done = now_s() - s->t0;
        if (r->c) emc_destroy(r->c);
        r->c = nullptr;
        s->used -= r->est_bytes;
        s->cv.notify_all();
      } else {
        r->state = EMC_REQ_RUNNING;
        s->active.push_back(r);
      }
    }
    std::vector<Req*> act;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      act = s->active;
    }
    if (act.empty()) continue;
    // group by mesh dimensions; a chunk ends on every member's window boundary
    std::map<std::vector<int>, std::vector<Req*>> groups;
    int chunk = 1 << 30;
    for (Req* r : act) {
      groups[{r->dims[0], r->dims[1], r->dims[2]}].push_back(r);
      chunk = std::min(chunk, r->o.window_steps - r->win_step);
    }
    std::map<std::vector<Req*>, emc_batch*> keep;
    std::vector<std::pair<emc_batch*, std::vector<Req*>>> work;
    for (auto& kv : groups) {
      auto it = batches.find(kv.second);
      emc_batch* b = nullptr;
      if (it != batches.end()) {
        b = it->second;
        batches.erase(it);
      } else {
        std::vector<emc_ctx*> cs;
        for (Req* r : kv.second) cs.push_back(r->c);
        int st = 0;
        b = emc_batch_create(cs.data(), (int)cs.size(), &st);
        if (st) {
          for (Req* r : kv.second) fail_req(r, st, emc_batch_last_error(b));
          emc_batch_destroy(b);
          b = nullptr;
        }
      }
      if (b) {
        keep[kv.second] = b;
        work.emplace_back(b, kv.second);
      }
    }
    drop_batches();  // groups that changed
    batches.swap(keep);
    // every group steps at once, each batch on its own stream
    std::vector<std::thread> th;
    std::vector<int> rc(work.size(), 0);
    std::vector<std::vector<emc_fly_tallies>> acc(work.size());
    for (size_t w = 0; w < work.size(); ++w) {
      acc[w].assign(work[w].second.size(), emc_fly_tallies{});
      th.emplace_back([&, w]() {
        (void)hipSetDevice(s->device);
        rc[w] = emc_batch_step(work[w].first, chunk, acc[w].data());
      });
    }
    for (auto& t : th) t.join();
    std::vector<Req*> done;
    std::lock_guard<std::mutex> lk(s->mu);  // results are read by emc_server_poll
    for (size_t w = 0; w < work.size(); ++w) {
      for (size_t q = 0; q < work[w].second.size(); ++q) {
        Req* r = work[w].second[q];
        if (rc[w]) {
          fail_req(r, rc[w], emc_batch_last_error(work[w].first));
          done.push_back(r);
          continue;
        }
        const emc_fly_tallies& t = acc[w][q];
        for (int i = 0; i < EMC_MAX_CONTACT; ++i)
          for (int j = 0; j < EMC_CT_NCOL; ++j) r->acc.contact[i][j] += t.contact[i][j];
        r->steps += chunk;
        r->win_step += chunk;
        if (r->win_step >= r->o.window_steps) {
          finish_window(r);
          (void)emc_tallies_reset(r->c);
        }
        if (r->converged || (int)r->win.size() >= r->o.max_windows) done.push_back(r);
      }
    }
    for (Req* r : act)
      if (r->status && std::find(done.begin(), done.end(), r) == done.end()) done.push_back(r);
    if (done.empty()) continue;
    drop_batches();
    for (Req* r : done) {
      emc_destroy(r->c);
      r->c = nullptr;
      s->used -= r->est_bytes;
      r->t_done = now_s() - s->t0;
      r->state = r->status ? EMC_REQ_FAILED : EMC_REQ_DONE;
      s->active.erase(std::find(s->active.begin(), s->active.end(), r));
    }
    s->cv.notify_all();
  }
  drop_batches();
}

}  // namespace

extern "C" {

void emc_default_request(emc_request_opts* r) {
  if (!r) return;
  r->seed = 1;
  r->window_steps = 1000;
  r->skip_windows = 1;
  r->min_windows = 4;
  r->max_windows = 64;
  r->rel_tol = 0.02;
  r->abs_tol = 0.0;
  r->drain_contact = 1;
  r->pad0 = 0;
}

emc_server* emc_server_create(const char* table_dir, int device, double hbm_budget_bytes, int max_active,
                              int* status) {
  int st = EMC_OK;
  emc_server* s = nullptr;
  int ndev = 0;
  if (!table_dir || hbm_budget_bytes <= 0 || max_active < 1) st = EMC_ERR_ARG;
  else if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) st = EMC_ERR_NOGPU;
  else {
    s = new emc_server();
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->table_dir = table_dir;
    emc_default_opts(&s->base);
    s->t0 = now_s();
    s->worker = std::thread(run_worker, s);
  }
  if (status) *status = st;
  return s;
}

int emc_server_submit(emc_server* s, const char* input_txt, const emc_request_opts* ro, int64_t* id) {
  if (!s || !input_txt) return EMC_ERR_ARG;
  emc_request_opts o;
  if (ro) o = *ro;
  else emc_default_request(&o);
  if (o.window_steps < 1 || o.max_windows < 1 || o.rel_tol < 0 || o.abs_tol < 0 || o.drain_contact >= EMC_MAX_CONTACT)
    return EMC_ERR_ARG;
  // the deck is read now, on the host (mesh and footprint), and the request queued
  emc_opts ho = s->base;
  ho.device = -1;
  ho.table_dir = s->table_dir.c_str();
  int st = 0;
  emc_ctx* c = emc_create(input_txt, &ho, &st);
  if (st) {
    s->err = emc_last_error(c);
    emc_destroy(c);
    return st;
  }
  emc_mesh_view m;
  emc_mesh_view_get(c, &m);
  std::unique_ptr<Req> r(new Req());
  r->deck = input_txt;
  r->o = o;
  r->dims[0] = m.cnx; r->dims[1] = m.cny; r->dims[2] = m.cnz;
  r->dt = m.dt;
  r->curr0 = m.curr0;
  r->nodes = m.nnode;
  r->est_bytes = estimate_bytes(m, r->deck, s->base.capacity_factor);
  emc_destroy(c);
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  r->t_submit = now_s() - s->t0;
  if (id) *id = r->id;
  s->queue.push_back(r.get());
  s->all[r->id] = std::move(r);
  s->cv.notify_all();
  return EMC_OK;
}

static void fill_result(const Req* r, emc_request_result* out) {
  std::memset(out, 0, sizeof(*out));
  out->state = r->state;
  out->status = r->status;
  out->id_mean = r->mean;
  out->id_half_width = r->half;
  out->converged = r->converged ? 1 : 0;
  out->windows = (int)r->win.size();
  for (int q = 0; q < out->windows && q < EMC_REQ_MAX_WINDOWS; ++q) out->id_windows[q] = r->win[q];
  out->steps = r->steps;
  out->particles = r->particles;
  out->nodes = r->nodes;
  out->footprint_bytes = r->bytes > 0 ? r->bytes : r->est_bytes;
  out->t_submit = r->t_submit;
  out->t_start = r->t_start;
  out->t_done = r->t_done;
  std::snprintf(out->error, sizeof(out->error), "%s", r->err.c_str());
}

int emc_server_poll(emc_server* s, int64_t id, emc_request_result* out) {
  if (!s || !out) return EMC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return EMC_ERR_ARG;
  fill_result(it->second.get(), out);
  return EMC_OK;
}

int emc_server_wait(emc_server* s, int64_t id, double timeout_s, emc_request_result* out) {
  if (!s || !out) return EMC_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return EMC_ERR_ARG;
  Req* r = it->second.get();
  auto finished = [&] { return r->state == EMC_REQ_DONE || r->state == EMC_REQ_FAILED; };
  if (timeout_s < 0) s->cv.wait(lk, finished);
  else s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
  fill_result(r, out);
  return finished() ? EMC_OK : EMC_ERR_STATE;
}

const char* emc_server_last_error(emc_server* s) { return s ? s->err.c_str() : "null server"; }

void emc_server_destroy(emc_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  for (auto& kv : s->all)
    if (kv.second->c) emc_destroy(kv.second->c);
  delete s;
}

}  // extern "C"
