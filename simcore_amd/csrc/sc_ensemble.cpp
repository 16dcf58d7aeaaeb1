// sc_ensemble.cpp -- ensemble runner: continuous batching of independent
// filament systems on one device.
//
// The reference runs the n_runs of a parameter file one after another
// (SimulationManager::RunSimulations, src/simulation_manager.cpp:461-545).
// Here requests arrive at any time and share the device: a request is
// admitted when a slot is free and its footprint fits the runner's arena,
// its state stays resident until its last step, and it is retired -- final
// positions copied out, arena blocks returned -- as soon as it finishes,
// which lets the next request in.
//
// Resident systems are stepped in chunks.  Systems that qualify
// (system_batchable) are stepped together: a batch is a set of up to
// batch_size systems whose chunk is one sequence of grouped launches
// (each kernel of the step covers every member, block ranges per member),
// so a step of sixteen small systems costs the launches of one.  Members
// join and leave at chunk boundaries; a chunk is an even number of steps,
// so every member starts it at an odd step.  The others run their chunks
// on their own streams (graph replay).  Systems never share buffers, and a
// member's blocks run the single-system kernels' code on its own data, so
// a result is the same whatever else runs beside it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "sc_internal.h"
#include "sc_system.h"

namespace sc {

int Arena::init(size_t bytes) {
  release();
  bytes = round(bytes);
  void *p = nullptr;
  if (bytes == 0 || hipMalloc(&p, bytes) != hipSuccess) return SC_ERR_NOMEM;
  base_ = static_cast<char *>(p);
  cap_ = bytes;
  free_.clear();
  free_[0] = bytes;
  taken_.clear();
  used_bytes_ = 0;
  return SC_OK;
}

void Arena::release() {
  if (base_) (void)hipFree(base_);
  base_ = nullptr;
  cap_ = 0;
  used_bytes_ = 0;
  free_.clear();
  taken_.clear();
}

void *Arena::get(size_t bytes) {
  bytes = round(bytes ? bytes : 1);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    const size_t off = it->first, sz = it->second;
    free_.erase(it);
    if (sz > bytes) free_[off + bytes] = sz - bytes;
    taken_[off] = bytes;
    used_bytes_ += bytes;
    return base_ + off;
  }
  return nullptr;
}

void Arena::put(void *p) {
  if (!p || !base_) return;
  const size_t off = static_cast<size_t>(static_cast<char *>(p) - base_);
  auto t = taken_.find(off);
  if (t == taken_.end()) return;
  const size_t sz = t->second;
  taken_.erase(t);
  used_bytes_ -= sz;
  auto it = free_.emplace(off, sz).first;
  auto nx = std::next(it);
  if (nx != free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_.erase(nx);
  }
  if (it != free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_.erase(it);
    }
  }
}

size_t Arena::largest_free() const {
  size_t m = 0;
  for (const auto &kv : free_) m = std::max(m, kv.second);
  return m;
}

}  // namespace sc

namespace {

struct Request {
  int64_t id = 0;
  sc_params p{};
  int32_t nf = 0;
  int64_t ns = 0;
  std::vector<int32_t> site_off;
  std::vector<double> r;
  std::vector<sc_filament_params> fil;
  int n_bins = 0;
  int64_t n_steps = 0;
  int64_t enqueued = 0;
  int32_t status = SC_REQ_QUEUED;
  double t_submit = 0, t_admit = -1, t_done = -1;
  size_t est_bytes = 0, bytes = 0;
  sc_handle h = nullptr;
  int slot = -1;
  int batch = -1;     // batch running its current chunk, -1: none
  bool solo = false;  // a chunk pending on the system's own stream
  std::vector<double> r_out;
  std::string err;
};

// One batch: its stream, member tables (pinned host copy, device copy),
// the control blocks gathered at the end of a chunk, and the members of
// the chunk in flight.
struct Batch {
  hipStream_t st = nullptr;
  sc::Member *hm = nullptr, *dm = nullptr;
  sc::Ctl *hctl = nullptr, *dctl = nullptr;
  hipEvent_t ev = nullptr;
  std::vector<int64_t> ids;
  bool busy = false;
  // an odd and an even step of the members (gx[0]) and kBatchGraphPairs
  // such pairs (gx[1]), replayed over a chunk; the launch grids depend only
  // on the members' sizes (key), the members' buffers are read from the
  // device table at run time
  hipGraphExec_t gx[2] = {nullptr, nullptr};
  std::vector<int64_t> gkey;
};

constexpr int kBatchGraphPairs = 4;

int count_bins(const std::vector<int32_t> &so) {
  const int bin_max[] = {8, 16, 32, 64, 128, 256};
  bool used[6] = {};
  for (size_t f = 0; f + 1 < so.size(); ++f) {
    const int n = so[f + 1] - so[f];
    for (int k = 0; k < 6; ++k)
      if (n <= bin_max[k]) {
        used[k] = true;
        break;
      }
  }
  int c = 0;
  for (bool u : used) c += u;
  return c;
}

}  // namespace

struct sc_ensemble_s {
  sc_ensemble_params prm{};
  sc::Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<int> free_slots;
  std::vector<std::unique_ptr<Request>> reqs;  // index = request id
  std::deque<int64_t> queue;
  std::vector<int64_t> running;
  std::vector<Batch> batches;
  sc::Ctl *ctl_pool = nullptr;  // pinned control-block mirror per slot
  std::chrono::steady_clock::time_point t0;
  std::string err;
  int32_t peak_active = 0;
  size_t peak_bytes = 0;
  int64_t batched_steps = 0, solo_steps = 0;
  double host_admit = 0, host_chunk = 0, host_retire = 0;  // seconds in each
  double host_capture = 0;  // of host_chunk: batch graph captures
  int64_t captures = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
};

namespace {

void retire(sc_ensemble e, Request &R) {
  if (R.h) sc_destroy(R.h);
  R.h = nullptr;
  if (R.slot >= 0) e->free_slots.push_back(R.slot);
  R.slot = -1;
  e->running.erase(std::remove(e->running.begin(), e->running.end(), R.id), e->running.end());
  // inputs are no longer needed once the request has left the device
  std::vector<double>().swap(R.r);
}

int enqueue_chunk(sc_ensemble e, Request &R) {
  const int64_t k = std::min<int64_t>(e->prm.steps_per_launch, R.n_steps - R.enqueued);
  if (k <= 0) return SC_OK;
  const int rc = sc_step_async(R.h, (int32_t)k);
  if (rc == SC_OK) {
    R.enqueued += k;
    R.solo = true;
    e->solo_steps += k;
  }
  return rc;
}

// Enqueues one chunk of batch B over the members in B.ids: k steps (even,
// at most every member's remaining steps), then the members' control
// blocks copied back, then B's event.
int enqueue_batch(sc_ensemble e, Batch &B) {
  const int n = (int)B.ids.size();
  // a chunk is at least one odd and one even step (every member has two left)
  int64_t k = std::max<int64_t>(2, e->prm.steps_per_launch);
  for (int64_t id : B.ids) k = std::min(k, e->reqs[id]->n_steps - e->reqs[id]->enqueued);
  k &= ~(int64_t)1;
  if (n == 0 || k <= 0) return SC_ERR_STATE;
  for (int j = 0; j < n; ++j) {
    Request &R = *e->reqs[B.ids[j]];
    sc::system_member(R.h, &B.hm[j]);
    R.batch = (int)(&B - e->batches.data());
    R.enqueued += k;
  }
  if (hipMemcpyAsync(B.dm, B.hm, sizeof(sc::Member) * n, hipMemcpyHostToDevice, B.st) !=
      hipSuccess)
    return SC_ERR_HIP;
  std::vector<int64_t> key;
  for (int j = 0; j < n; ++j) {
    const sc::Geo &g = B.hm[j].g;
    key.insert(key.end(), {g.nb, g.ns, (int64_t)g.n1d, (int64_t)B.hm[j].pack_waves});
  }
  if (!B.gx[0] || key != B.gkey) {
    const double tcap = e->now();
    ++e->captures;
    B.gkey.clear();
    for (int v = 0; v < 2; ++v) {
      if (B.gx[v]) (void)hipGraphExecDestroy(B.gx[v]);
      B.gx[v] = nullptr;
    }
    for (int v = 0; v < 2; ++v) {
      if (hipStreamBeginCapture(B.st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return SC_ERR_HIP;
      for (int q = 0; q < (v ? kBatchGraphPairs : 1); ++q) {
        // odd step (the rebuild stage, then the midstep), even step
        sc::launch_group_rebuild(B.dm, B.hm, n, B.st);
        sc::launch_group_pair(B.dm, B.hm, n, true, B.st);
        sc::launch_group_integrate(B.dm, B.hm, n, true, B.st);
        sc::launch_group_commit(B.dm, n, B.st);
        sc::launch_group_pair(B.dm, B.hm, n, false, B.st);
        sc::launch_group_integrate(B.dm, B.hm, n, false, B.st);
        sc::launch_group_commit(B.dm, n, B.st);
      }
      hipGraph_t gr = nullptr;
      if (hipStreamEndCapture(B.st, &gr) != hipSuccess) return SC_ERR_HIP;
      const hipError_t ie = hipGraphInstantiate(&B.gx[v], gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (ie != hipSuccess) {
        B.gx[v] = nullptr;
        return SC_ERR_HIP;
      }
    }
    B.gkey.swap(key);
    e->host_capture += e->now() - tcap;
  }
  for (int64_t s = 0; s < k;) {
    const bool big = k - s >= 2 * kBatchGraphPairs;
    if (hipGraphLaunch(B.gx[big ? 1 : 0], B.st) != hipSuccess) return SC_ERR_HIP;
    s += big ? 2 * kBatchGraphPairs : 2;
  }
  sc::launch_group_ctl(B.dm, n, B.dctl, B.st);
  if (hipMemcpyAsync(B.hctl, B.dctl, sizeof(sc::Ctl) * n, hipMemcpyDeviceToHost, B.st) !=
          hipSuccess ||
      hipGetLastError() != hipSuccess || hipEventRecord(B.ev, B.st) != hipSuccess)
    return SC_ERR_HIP;
  B.busy = true;
  e->batched_steps += k * n;
  return SC_OK;
}

void fail_request(sc_ensemble e, Request &R, int rc) {
  R.status = rc;
  if (R.err.empty()) R.err = R.h ? sc_last_error(R.h) : "request failed";
  R.t_done = e->now();
  retire(e, R);
}

// R's last chunk has ended (R idle): retire it if finished
bool finish_if_done(sc_ensemble e, Request &R) {
  if (R.enqueued < R.n_steps) return false;
  (void)sc_device_bytes(R.h, &R.bytes);  // the footprint it ended with
  R.r_out.resize(3 * (size_t)R.ns);
  const int rc = sc_get_sites(R.h, R.r_out.data(), nullptr, nullptr);
  if (rc != SC_OK) {
    fail_request(e, R, rc);
    return true;
  }
  R.status = SC_REQ_DONE;
  R.t_done = e->now();
  retire(e, R);
  return true;
}

// Creates and loads R's system in slot `slot`; SC_ERR_NOMEM leaves R queued.
int admit(sc_ensemble e, Request &R, int slot) {
  sc_handle h = nullptr;
  int rc = sc::system_create(&R.p, e->streams[slot], &e->arena, &h, &e->ctl_pool[slot]);
  if (rc) return rc;
  rc = sc_load_filaments(h, R.nf, R.site_off.data(), R.r.data(), R.fil.data());
  if (rc) {
    if (rc != SC_ERR_NOMEM) R.err = sc_last_error(h);
    sc_destroy(h);
    return rc;
  }
  R.h = h;
  R.slot = slot;
  (void)sc_device_bytes(h, &R.bytes);
  R.status = SC_REQ_RUNNING;
  R.t_admit = e->now();
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_ensemble_create(const sc_ensemble_params *params, sc_ensemble *out) {
  if (!out) return SC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SC_ERR_HIP;
  std::unique_ptr<sc_ensemble_s> e(new sc_ensemble_s());
  if (params) e->prm = *params;
  if (e->prm.max_active <= 0) e->prm.max_active = 16;
  if (e->prm.steps_per_launch <= 0) e->prm.steps_per_launch = 64;
  size_t budget = e->prm.hbm_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return SC_ERR_HIP;
    budget = (size_t)(0.8 * (double)fr);
  }
  e->prm.hbm_budget = budget;
  if (e->arena.init(budget) != SC_OK) return SC_ERR_NOMEM;
  {
    void *pin = nullptr;
    if (hipHostMalloc(&pin, sizeof(sc::Ctl) * e->prm.max_active, hipHostMallocDefault) !=
        hipSuccess)
      return SC_ERR_HIP;
    e->ctl_pool = static_cast<sc::Ctl *>(pin);
  }
  for (int k = 0; k < e->prm.max_active; ++k) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t s : e->streams) (void)hipStreamDestroy(s);
      (void)hipHostFree(e->ctl_pool);
      return SC_ERR_HIP;
    }
    e->streams.push_back(st);
    e->free_slots.push_back(e->prm.max_active - 1 - k);
  }
  if (e->prm.batch_size <= 0) e->prm.batch_size = 8;
  e->prm.batch_size = std::min<int32_t>(e->prm.batch_size, sc::kMaxGroup);
  if (e->prm.batch_size > 1) {
    // enough batches for every resident system, at most four (one chunk
    // runs while the host sets up the next)
    const int nbat =
        std::min(4, (e->prm.max_active + e->prm.batch_size - 1) / e->prm.batch_size);
    e->batches.resize(nbat);
    for (Batch &B : e->batches) {
      void *hm = nullptr, *hc = nullptr;
      if (hipStreamCreateWithFlags(&B.st, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&B.ev, hipEventDisableTiming) != hipSuccess ||
          hipHostMalloc(&hm, sizeof(sc::Member) * sc::kMaxGroup, hipHostMallocDefault) !=
              hipSuccess ||
          hipHostMalloc(&hc, sizeof(sc::Ctl) * sc::kMaxGroup, hipHostMallocDefault) !=
              hipSuccess ||
          hipMalloc(&B.dm, sizeof(sc::Member) * sc::kMaxGroup) != hipSuccess ||
          hipMalloc(&B.dctl, sizeof(sc::Ctl) * sc::kMaxGroup) != hipSuccess) {
        B.hm = static_cast<sc::Member *>(hm);
        B.hctl = static_cast<sc::Ctl *>(hc);
        sc_ensemble_destroy(e.release());
        return SC_ERR_HIP;
      }
      B.hm = static_cast<sc::Member *>(hm);
      B.hctl = static_cast<sc::Ctl *>(hc);
    }
  }
  e->t0 = std::chrono::steady_clock::now();
  *out = e.release();
  return SC_OK;
}

int sc_ensemble_submit(sc_ensemble e, const sc_params *params, int32_t n_filaments,
                       const int32_t *site_off, const double *r, const sc_filament_params *fil,
                       int64_t n_steps, double t_submit, int64_t *request_id) {
  if (!e) return SC_ERR_ARG;
  if (!params || !site_off || !r || !fil || n_filaments <= 0 || n_steps < 0)
    return e->fail(SC_ERR_ARG, "bad request");
  std::unique_ptr<Request> R(new Request());
  R->id = (int64_t)e->reqs.size();
  R->p = *params;
  R->nf = n_filaments;
  R->site_off.assign(site_off, site_off + n_filaments + 1);
  R->ns = R->site_off.back();
  if (R->site_off[0] != 0 || R->ns <= 0) return e->fail(SC_ERR_ARG, "bad site offsets");
  R->r.assign(r, r + 3 * R->ns);
  R->fil.assign(fil, fil + n_filaments);
  R->n_bins = count_bins(R->site_off);
  R->n_steps = n_steps;
  R->t_submit = t_submit < 0 ? e->now() : t_submit;
  R->est_bytes = sc::system_estimate_bytes(&R->p, R->nf, R->ns, R->n_bins);
  if (request_id) *request_id = R->id;
  if (R->est_bytes > e->arena.capacity()) {
    R->status = SC_ERR_NOMEM;
    R->err = "request footprint exceeds the runner's device budget";
    R->t_done = R->t_submit;
  } else {
    e->queue.push_back(R->id);
  }
  e->reqs.push_back(std::move(R));
  return SC_OK;
}

int sc_ensemble_progress(sc_ensemble e, int32_t *n_running, int32_t *n_queued) {
  if (!e) return SC_ERR_ARG;
  // 1. batches whose chunk has ended: each member's outcome as its own
  //    poll would report it
  for (Batch &B : e->batches) {
    if (!B.busy) continue;
    const hipError_t q = hipEventQuery(B.ev);
    if (q == hipErrorNotReady) continue;
    B.busy = false;
    for (size_t j = 0; j < B.ids.size(); ++j) {
      Request &R = *e->reqs[B.ids[j]];
      R.batch = -1;
      int done = 0;
      const int rc = q != hipSuccess ? SC_ERR_HIP
                                     : sc::system_batch_done(R.h, B.hctl[j], R.enqueued, &done);
      if (rc != SC_OK) {
        if (q != hipSuccess) R.err = hipGetErrorString(q);
        fail_request(e, R, rc);
      } else if (!done) {
        R.solo = true;  // a capacity re-run on the system's own stream
      }
    }
    B.ids.clear();
  }
  // 2. chunks on the systems' own streams
  const std::vector<int64_t> run = e->running;
  for (int64_t id : run) {
    Request &R = *e->reqs[id];
    if (!R.solo) continue;
    int done = 0;
    const int rc = sc_step_poll(R.h, &done);
    if (rc != SC_OK) {
      fail_request(e, R, rc);
      continue;
    }
    if (done) R.solo = false;
  }
  // 3. finished systems leave (before admission, which may use their room)
  const std::vector<int64_t> idle_scan = e->running;
  const double tr = e->now();
  for (int64_t id : idle_scan) {
    Request &R = *e->reqs[id];
    if (!R.solo && R.batch < 0) finish_if_done(e, R);
  }
  e->host_retire += e->now() - tr;
  // 4. admission, first fit in arrival order
  for (auto it = e->queue.begin(); it != e->queue.end() && !e->free_slots.empty();) {
    Request &R = *e->reqs[*it];
    if (R.est_bytes > e->arena.largest_free()) {
      ++it;
      continue;
    }
    const int slot = e->free_slots.back();
    const double ta = e->now();
    const int rc = admit(e, R, slot);
    e->host_admit += e->now() - ta;
    if (rc == SC_ERR_NOMEM) {
      ++it;
      continue;
    }
    e->free_slots.pop_back();
    it = e->queue.erase(it);
    if (rc != SC_OK) {
      R.status = rc;
      if (R.err.empty()) R.err = "admission failed";
      R.t_done = e->now();
      if (R.h) sc_destroy(R.h);
      R.h = nullptr;
      e->free_slots.push_back(slot);
      continue;
    }
    e->running.push_back(R.id);
    e->peak_active = std::max<int32_t>(e->peak_active, (int32_t)e->running.size());
    e->peak_bytes = std::max(e->peak_bytes, e->arena.used());
  }
  // 5. idle systems take their next chunk, in a free batch when they
  //    qualify, else on their own stream (a request of 0 steps finishes)
  const double tc = e->now();
  const std::vector<int64_t> idle_now = e->running;
  std::vector<int64_t> ready;
  for (int64_t id : idle_now) {
    Request &R = *e->reqs[id];
    if (R.solo || R.batch >= 0) continue;
    if (finish_if_done(e, R)) continue;
    if (!e->batches.empty() && R.n_steps - R.enqueued >= 2 && sc::system_batchable(R.h)) {
      ready.push_back(id);
      continue;
    }
    const int rc = enqueue_chunk(e, R);
    if (rc != SC_OK) fail_request(e, R, rc);
  }
  if (!ready.empty()) {
    // largest first onto the free batch with the fewest bonds so far
    std::sort(ready.begin(), ready.end(), [&](int64_t a, int64_t b) {
      return e->reqs[a]->ns - e->reqs[a]->nf > e->reqs[b]->ns - e->reqs[b]->nf;
    });
    std::vector<int64_t> load(e->batches.size(), 0);
    for (int64_t id : ready) {
      int best = -1;
      for (size_t b = 0; b < e->batches.size(); ++b) {
        const Batch &B = e->batches[b];
        if (B.busy || (int)B.ids.size() >= e->prm.batch_size) continue;
        if (best < 0 || load[b] < load[best]) best = (int)b;
      }
      if (best < 0) break;  // every batch busy or full: wait for a boundary
      e->batches[best].ids.push_back(id);
      load[best] += e->reqs[id]->ns - e->reqs[id]->nf;
    }
    for (Batch &B : e->batches) {
      if (B.busy || B.ids.empty()) continue;
      const int rc = enqueue_batch(e, B);
      if (rc != SC_OK) {
        // nothing of the chunk is known to have run: fail its members
        for (int64_t id : B.ids) {
          Request &R = *e->reqs[id];
          R.batch = -1;
          R.err = "batch launch failed";
          fail_request(e, R, rc);
        }
        B.ids.clear();
        B.busy = false;
      }
    }
  }
  e->host_chunk += e->now() - tc;
  if (n_running) *n_running = (int32_t)e->running.size();
  if (n_queued) *n_queued = (int32_t)e->queue.size();
  return SC_OK;
}

int sc_ensemble_drain(sc_ensemble e, double timeout_s) {
  if (!e) return SC_ERR_ARG;
  const double t_end = timeout_s > 0 ? e->now() + timeout_s : 0.0;
  for (;;) {
    int32_t nr = 0, nq = 0;
    int rc = sc_ensemble_progress(e, &nr, &nq);
    if (rc) return rc;
    if (nr == 0 && nq == 0) return SC_OK;
    if (nr == 0 && nq > 0) {
      // nothing resident and the queue head does not fit an empty arena:
      // only possible through fragmentation left by failed admissions
      return e->fail(SC_ERR_NOMEM, "queued requests cannot be admitted");
    }
    if (timeout_s > 0 && e->now() > t_end) return e->fail(SC_ERR_STATE, "drain timed out");
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

int sc_ensemble_request(sc_ensemble e, int64_t id, sc_request_info *info) {
  if (!e || !info) return SC_ERR_ARG;
  if (id < 0 || id >= (int64_t)e->reqs.size()) return e->fail(SC_ERR_ARG, "unknown request");
  const Request &R = *e->reqs[id];
  info->status = R.status;
  info->n_filaments = R.nf;
  info->n_sites = R.ns;
  info->n_bonds = R.ns - R.nf;
  info->n_steps = R.n_steps;
  // while running: steps enqueued so far (reading the device counters
  // would wait for the handle's stream)
  info->steps_done = R.status == SC_REQ_DONE ? R.n_steps
                     : R.status == SC_REQ_RUNNING ? R.enqueued
                                                  : 0;
  info->t_submit = R.t_submit;
  info->t_admit = R.t_admit;
  info->t_done = R.t_done;
  info->device_bytes = R.bytes;
  return SC_OK;
}

int sc_ensemble_result(sc_ensemble e, int64_t id, double *r) {
  if (!e || !r) return SC_ERR_ARG;
  if (id < 0 || id >= (int64_t)e->reqs.size()) return e->fail(SC_ERR_ARG, "unknown request");
  const Request &R = *e->reqs[id];
  if (R.status < 0) return e->fail(R.status, "request failed: " + R.err);
  if (R.status != SC_REQ_DONE) return e->fail(SC_ERR_STATE, "request not finished");
  std::copy(R.r_out.begin(), R.r_out.end(), r);
  return SC_OK;
}

int sc_ensemble_stats_get(sc_ensemble e, sc_ensemble_stats *out) {
  if (!e || !out) return SC_ERR_ARG;
  *out = sc_ensemble_stats{};
  std::vector<double> lat;
  double first = -1, last = -1;
  for (const auto &p : e->reqs) {
    const Request &R = *p;
    ++out->submitted;
    if (R.status == SC_REQ_DONE) {
      ++out->completed;
      lat.push_back(R.t_done - R.t_submit);
      out->segment_steps += (double)(R.ns - R.nf) * (double)R.n_steps;
      last = std::max(last, R.t_done);
    } else if (R.status < 0) {
      ++out->failed;
    } else if (R.status == SC_REQ_RUNNING) {
      ++out->running;
    } else {
      ++out->queued;
    }
    first = first < 0 ? R.t_submit : std::min(first, R.t_submit);
  }
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto rank = [&](double q) {
      // nearest-rank percentile
      size_t k = (size_t)std::ceil(q * (double)lat.size());
      if (k < 1) k = 1;
      return lat[std::min(k, lat.size()) - 1];
    };
    out->latency_p50 = rank(0.50);
    out->latency_p99 = rank(0.99);
    double s = 0;
    for (double v : lat) s += v;
    out->latency_mean = s / (double)lat.size();
    out->latency_max = lat.back();
    out->span = last - first;
  }
  out->peak_active = e->peak_active;
  out->peak_bytes = e->peak_bytes;
  out->arena_bytes = e->arena.capacity();
  out->batched_steps = e->batched_steps;
  out->solo_steps = e->solo_steps;
  out->host_admit_s = e->host_admit;
  out->host_chunk_s = e->host_chunk;
  out->host_retire_s = e->host_retire;
  out->host_capture_s = e->host_capture;
  out->graph_captures = e->captures;
  return SC_OK;
}

double sc_ensemble_clock(sc_ensemble e) { return e ? e->now() : 0.0; }

const char *sc_ensemble_last_error(sc_ensemble e) {
  return e ? e->err.c_str() : "null ensemble";
}

int sc_ensemble_destroy(sc_ensemble e) {
  if (!e) return SC_ERR_ARG;
  for (Batch &B : e->batches) {
    if (B.st) (void)hipStreamSynchronize(B.st);
    for (hipGraphExec_t g : B.gx)
      if (g) (void)hipGraphExecDestroy(g);
    if (B.dm) (void)hipFree(B.dm);
    if (B.dctl) (void)hipFree(B.dctl);
    if (B.hm) (void)hipHostFree(B.hm);
    if (B.hctl) (void)hipHostFree(B.hctl);
    if (B.ev) (void)hipEventDestroy(B.ev);
    if (B.st) (void)hipStreamDestroy(B.st);
  }
  e->batches.clear();
  for (auto &p : e->reqs)
    if (p->h) {
      sc_destroy(p->h);
      p->h = nullptr;
    }
  for (hipStream_t s : e->streams) (void)hipStreamDestroy(s);
  if (e->ctl_pool) (void)hipHostFree(e->ctl_pool);
  e->arena.release();
  delete e;
  return SC_OK;
}

}  // extern "C"
