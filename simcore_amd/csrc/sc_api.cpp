// sc_api.cpp -- C ABI of the device path: handle lifetime, HBM allocation,
// step orchestration (eager or hipGraph), candidate-capacity growth, taps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sc_internal.h"
#include "sc_math.h"
#include "sc_system.h"

using sc::Ctl;
using sc::Dev;
using sc::Geo;

namespace {

struct Bin {
  int maxn = 0;
  int count = 0;
  int32_t *list = nullptr;
  std::vector<int64_t> last;  // last bond of each listed filament (ascending)
};

constexpr int kFamilies = 4;  // pair+scatter, commit, integrate, rebuild

}  // namespace

struct sc_system {
  sc_params p{};
  Geo g{};
  Dev d{};
  hipStream_t stream = nullptr;
  std::string err;
  bool loaded = false;
  std::vector<int32_t> site_off;  // storage order
  // spatial order: storage filament f holds the caller's filament perm[f];
  // site/bond maps send storage indices to the caller's
  bool reordered = false;
  std::vector<int32_t> perm;
  std::vector<int64_t> site_map, bond_map;
  std::vector<Bin> bins;
  // packed integrator layout (mixed precision): lane table and per-wave
  // lane-move class of the filaments of at most 128 sites
  int32_t *pack_tab = nullptr;
  int pack_cs[4] = {0, 0, 0, 0};  // waves of lane-group width 16 << c: [pack_cs[c], pack_cs[c+1])
  int pack_waves = 0;
  int64_t pack_sites = 0;
  sc::Tune tune;
  // graphs: [0] odd step (rebuild stage included), [1] even step (a rebuild
  // is only ever decided after a full step, i.e. at odd steps), [2] the
  // odd+even pair, [3] kGraphPairs such pairs (fewer graph boundaries on
  // long runs)
  static constexpr int kGraphs = 4;
  static constexpr int kGraphPairs = 4;
  hipGraphExec_t gexec[kGraphs] = {};
  hipGraph_t graph[kGraphs] = {};
  // profiling
  bool prof = false;
  hipEvent_t ev[2 * kFamilies] = {};
  double prof_ms[kFamilies] = {};
  int64_t prof_n[kFamilies] = {};
  // asynchronous stepping: work enqueued up to step index `target`; the
  // pinned Ctl mirror is refreshed at the end of every enqueued range and
  // `done_ev` marks the end of the last one
  int64_t i_step_host = 0;  // i_step when nothing is pending
  // the two bond-record buffers: the state after i completed steps is in
  // rec[i & 1] (each step reads one and writes the other)
  double *rec[2] = {nullptr, nullptr};
  int64_t target = 0;
  bool pending = false;
  int grow_rounds = 0;
  hipEvent_t done_ev = nullptr;
  Ctl *ctl_host = nullptr;
  // length bins spread over side streams (fork/join around the integrator,
  // captured into the graphs as parallel branches)
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kSide] = {};
  // overlapped pair pass / integrator: one event per bond range
  hipEvent_t part_ev[sc::kMaxParts] = {};
  int64_t long_last = -1;  // last bond of the filaments above 128 sites (-1: none)
  std::vector<int64_t> pack_last;  // per packed wave, its filaments' last bond
                                   // (ascending inside each width group)
  // allocations (from `arena` when set, else hipMalloc)
  sc::Arena *arena = nullptr;
  bool own_stream = true;
  bool own_ctl_host = true;  // ctl_host allocated here (else the caller's)
  int device = 0;  // HIP device current at creation
  double *xi_dev = nullptr;  // caller-supplied noise (sc_set_noise)
  double max_bond_length = 0.0;
  std::vector<void *> fixed_allocs;
  std::vector<void *> cand_allocs;
  size_t fixed_bytes = 0, cand_bytes = 0;
};

namespace {

// Scratch of sc_theta_histogram_noalloc: histogram counters in a file-scope
// device array, so that entry point allocates nothing.  Each GPU holds its
// own copy for the life of the process; its device address is looked up
// once per device (scratch_setup, from system_create) and used from then on
// -- the array's name is a host-side stand-in, never a device address.
constexpr int64_t kScratchCounters = (int64_t)1 << 16;  // 512 KiB
static __device__ __align__(256) unsigned long long scratch[kScratchCounters];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: the device address of its scratch copy.
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<unsigned long long *>(ptr);
  return hipSuccess;
}

unsigned long long *scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// q-th cell of the half shell of cell k (CellList::AssignCellNeighbors
// order, src/cell_list.cpp:206-262), host side.
int32_t host_shell_key(int n, int32_t k, int q) {
  const int x = k / (n * n), y = (k / n) % n, z = k % n;
  int dx = 1, dy = 0, dz = 0;
  if (q >= 1 && q < 4) {
    dx = q - 2;
    dy = 1;
  } else if (q >= 4) {
    dx = (q - 4) % 3 - 1;
    dy = (q - 4) / 3 - 1;
    dz = 1;
  }
  const int xp = (x + dx + n) % n, yp = (y + dy + n) % n, zp = (z + dz) % n;
  return ((int32_t)xp * n + yp) * n + zp;
}

// Geo::home_order: for each class of cell position (on the low face, in
// the interior or on the high face in x and y; on the low face in z or
// not), the 14 home cells of a bond -- its own cell (code 13) and the 13
// cells whose half shell holds it (code q: half-shell offset q negated) --
// sorted by cell key.  With at least 3 cells per side the order depends on
// the class only: keys compare as (x, y, z) lexicographically and each
// coordinate's wrapped offsets keep one order inside a class.
void build_home_order(sc::Geo &g) {
  const int n = g.n1d;
  const int xs[3] = {0, 1, n - 1};
  for (int cxc = 0; cxc < 3; ++cxc)
    for (int cyc = 0; cyc < 3; ++cyc)
      for (int czc = 0; czc < 2; ++czc) {
        const int x = xs[cxc], y = xs[cyc], z = czc == 0 ? 0 : 1;
        const int32_t k = ((int32_t)x * n + y) * n + z;
        std::vector<std::pair<int32_t, int>> cells;
        cells.push_back({k, 13});
        for (int q = 0; q < 13; ++q) {
          // the mirror of half-shell cell q: offsets negated
          const int32_t fk = host_shell_key(n, k, q);
          const int fx = fk / (n * n), fy = (fk / n) % n, fz = fk % n;
          const int mx = (2 * x - fx + 2 * n) % n, my = (2 * y - fy + 2 * n) % n,
                    mz = (2 * z - fz + 2 * n) % n;
          cells.push_back({((int32_t)mx * n + my) * n + mz, q});
        }
        std::sort(cells.begin(), cells.end());
        uint64_t code = 0;
        for (int i = 13; i >= 0; --i) code = (code << 4) | (uint64_t)cells[i].second;
        g.home_order[cxc * 6 + cyc * 2 + czc] = code;
      }
}

// InteractionManager::InitInteractions -> CellList::SetMinCellLength /
// Init (src/interaction_manager.cpp:46-54, src/cell_list.cpp:9-64)
int cells_per_side(const sc_params &p) {
  const double rcut = std::pow(2.0, 1.0 / 6.0) * p.wca_sig;
  double min_len = p.min_cell_length;
  const double rc_len = std::sqrt(rcut * rcut);
  if (rc_len > min_len) min_len = rc_len;
  int n1d = (int)std::floor(2 * p.system_radius / min_len);
  if (n1d <= 3) n1d = 3;
  return n1d;
}

int fail(sc_system *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, expr)                                                       \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return fail((h), SC_ERR_HIP,                                            \
                  std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

void dfree(sc_system *h, void *v) {
  if (h->arena)
    h->arena->put(v);
  else
    (void)hipFree(v);
}

template <typename T>
int dalloc(sc_system *h, T **p, size_t n, bool cand) {
  size_t bytes = sc::Arena::round(sizeof(T) * (n ? n : 1));
  void *v = nullptr;
  if (h->arena) {
    v = h->arena->get(bytes);
    if (!v) {
      *p = nullptr;
      return fail(h, SC_ERR_NOMEM, "arena exhausted (" + std::to_string(bytes) + " bytes)");
    }
  } else {
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(h, SC_ERR_NOMEM,
                  "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
  }
  *p = static_cast<T *>(v);
  if (cand) {
    h->cand_allocs.push_back(v);
    h->cand_bytes += bytes;
  } else {
    h->fixed_allocs.push_back(v);
    h->fixed_bytes += bytes;
  }
  return SC_OK;
}

void drop_graph(sc_system *h) {
  for (int k = 0; k < sc_system::kGraphs; ++k) {
    if (h->gexec[k]) (void)hipGraphExecDestroy(h->gexec[k]);
    if (h->graph[k]) (void)hipGraphDestroy(h->graph[k]);
    h->gexec[k] = nullptr;
    h->graph[k] = nullptr;
  }
}

void free_cand(sc_system *h) {
  for (void *v : h->cand_allocs) dfree(h, v);
  h->cand_allocs.clear();
  h->cand_bytes = 0;
}

// Candidates are numbered in int32 and each incidence entry holds a
// candidate index shifted left by one bit.
constexpr int64_t kMaxCandidates = ((int64_t)1 << 30) - 1;

// Initial candidate capacity: capacity_factor * bonds + 4096 (4 when the
// factor is 0); a negative factor is an absolute capacity of -factor
// candidates (tests of capacity growth use it).
int64_t initial_capacity(const sc_params &p, int64_t nb) {
  if (p.capacity_factor < 0) return std::max<int64_t>(1, (int64_t)(-p.capacity_factor));
  const double cf = p.capacity_factor > 0 ? p.capacity_factor : 4.0;
  return std::min<int64_t>((int64_t)(cf * nb) + 4096, kMaxCandidates);
}

int alloc_cand(sc_system *h, int64_t cap) {
  if (cap > kMaxCandidates)
    return fail(h, SC_ERR_NOMEM, "more than 2^30 - 1 candidate pairs (" + std::to_string(cap) +
                                     " asked for)");
  free_cand(h);
  drop_graph(h);
  Dev &d = h->d;
  d.inc = nullptr;
  d.islot64 = nullptr;
  d.islot32 = nullptr;
  int rc;
  if ((rc = dalloc(h, &d.inc, 2 * cap, true))) return rc;
  if (h->p.precision == SC_PREC_FP64) {
    if ((rc = dalloc(h, &d.islot64, 2 * cap, true))) return rc;
  } else {
    if ((rc = dalloc(h, &d.islot32, 2 * cap, true))) return rc;
  }
  h->g.cap = cap;
  return SC_OK;
}

// Capacity after an overflow that asked for `need` candidates: doubled, or
// need + 25% + 1024 if more, at most kMaxCandidates unless need itself is
// larger (alloc_cand then refuses).
int64_t grown_capacity(int64_t cap, int64_t need) {
  const int64_t want = std::max<int64_t>(2 * cap, need + need / 4 + 1024);
  return need > kMaxCandidates ? need : std::min(want, kMaxCandidates);
}

// Control-block transfers, ordered on the system's stream.
Ctl read_ctl(sc_system *h) {
  Ctl c{};
  (void)hipMemcpyAsync(&c, h->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream);
  (void)hipStreamSynchronize(h->stream);
  return c;
}

int write_ctl(sc_system *h, const Ctl &c) {
  HIPCHK(h, hipMemcpyAsync(h->d.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return SC_OK;
}

// Points d.br / d.br_out at the records of the state after i_step steps.
// Host code that reads the records or launches kernels on the current
// state calls it once the stream is idle (ctl_host then holds i_step).
void refresh_records(sc_system *h) {
  if (!h->rec[0]) return;
  double *cur = h->rec[h->ctl_host->i_step & 1];
  h->d.br = cur;
  h->d.br_out = cur;
}

// The side streams and join events of the integrator's length bins, made
// the first time a launch shape asks for them (most systems never do, and
// an ensemble admits many systems).
int ensure_side_streams(sc_system *h) {
  if (h->side[0]) return SC_OK;
  for (int k = 0; k < sc_system::kSide; ++k)
    if (hipStreamCreateWithFlags(&h->side[k], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->join_ev[k], hipEventDisableTiming) != hipSuccess)
      return fail(h, SC_ERR_HIP, "side stream creation failed");
  for (int k = 0; k < sc::kMaxParts; ++k)
    if (hipEventCreateWithFlags(&h->part_ev[k], hipEventDisableTiming) != hipSuccess)
      return fail(h, SC_ERR_HIP, "event creation failed");
  return SC_OK;
}

// Whether the midsteps among steps from+1 .. from+n (midstep index
// (k - 1) / 2 of odd step k) all lie in the caller's noise buffer.
bool midsteps_in_buffer(const sc_system *h, int64_t from, int64_t n) {
  int64_t k0 = from + 1, k1 = from + n;
  if (!(k0 & 1)) ++k0;
  if (!(k1 & 1)) --k1;
  if (k1 < k0) return true;  // no midstep
  return (k0 - 1) / 2 >= h->g.xi_m0 && (k1 - 1) / 2 < h->g.xi_m0 + h->g.xi_nm;
}

// One step.  `odd` says whether the step index is odd: only odd steps can
// rebuild, so even steps skip the rebuild stage's launches.
void enqueue_step(sc_system *h, bool odd) {
  // an odd step k starts from the state after k - 1 (even) steps
  Dev d = h->d;
  d.br = h->rec[odd ? 0 : 1];
  d.br_out = h->rec[odd ? 1 : 0];
  const Geo &g = h->g;
  hipStream_t st = h->stream;
  const bool pf = h->prof;
  if (pf) (void)hipEventRecord(h->ev[6], st);
  if (odd) sc::launch_rebuild(d, g, st);
  if (pf) {
    (void)hipEventRecord(h->ev[7], st);
    (void)hipEventRecord(h->ev[0], st);
  }
  // Overlapped pair pass and integrator (no crosslinks, not profiling): the
  // bonds are cut into P ranges; the pair pass runs range by range on the
  // system's stream, and the integrator of the filaments whose bonds all
  // lie below a cut runs on a side stream as soon as the ranges below it
  // are done, beside the pair pass of the next range.  Each step reads one
  // record buffer and writes the other, so the two never touch the same
  // memory; the results are those of the sequential order.
  // Early integration of the long filaments (no crosslinks, not
  // profiling): with the filaments above 128 sites stored first, the pair
  // pass is cut after their last bond; their integrator -- few wavefronts,
  // bound by latency -- then runs on a side stream beside the rest of the
  // pair pass, and the packed filaments follow the pass on the system's
  // stream.  Same memory argument as below; the results are those of the
  // sequential order.
  if (h->tune.early_long && g.nx == 0 && h->side[1] && !pf && h->long_last >= 0) {
    const int64_t nb = g.nb;
    const int64_t cut = std::min<int64_t>(nb, (h->long_last / sc::kPairAlign + 1) * sc::kPairAlign);
    hipStream_t sl = h->side[1];
    sc::launch_pair_range(d, g, h->p.precision, h->tune, st, g.thermo != 0, 0, cut);
    (void)hipEventRecord(h->part_ev[0], st);
    (void)hipStreamWaitEvent(sl, h->part_ev[0], 0);
    const bool packed = h->tune.int_layout >= 2 && h->pack_waves > 0;
    for (const Bin &b : h->bins)
      if (b.maxn > 128)
        sc::launch_integrate(d, g, b.list, b.count, b.maxn, h->p.precision, h->tune, odd, sl);
    (void)hipEventRecord(h->join_ev[1], sl);
    sc::launch_pair_range(d, g, h->p.precision, h->tune, st, g.thermo != 0, cut, nb);
    if (packed && h->tune.int_layout == 2) {
      sc::launch_integrate_packed_all(d, g, h->pack_tab, h->pack_waves, h->pack_cs[1],
                                      h->pack_cs[2], h->tune.int_waves, odd, st);
    } else if (packed) {
      for (int c = 0; c < 3; ++c)
        if (h->pack_cs[c + 1] > h->pack_cs[c])
          sc::launch_integrate_packed(d, g, h->pack_tab + (int64_t)64 * h->pack_cs[c],
                                      h->pack_cs[c + 1] - h->pack_cs[c], 16 << c,
                                      h->tune.int_waves, odd, st);
    }
    for (const Bin &b : h->bins)
      if (b.maxn <= 128 && !packed)
        sc::launch_integrate(d, g, b.list, b.count, b.maxn, h->p.precision, h->tune, odd, st);
    (void)hipStreamWaitEvent(st, h->join_ev[1], 0);
    sc::launch_commit(d, g, st);
    return;
  }
  const int P = std::min(h->tune.parts, sc::kMaxParts);
  if (P > 1 && g.nx == 0 && h->side[0] && !pf) {
    const int64_t nb = g.nb;
    int64_t cut[sc::kMaxParts + 1];
    cut[0] = 0;
    for (int p = 1; p < P; ++p) cut[p] = std::max(cut[p - 1], nb * p / P / sc::kPairAlign * sc::kPairAlign);
    cut[P] = nb;
    hipStream_t si = h->side[0];
    for (int p = 0; p < P; ++p) {
      sc::launch_pair_range(d, g, h->p.precision, h->tune, st, g.thermo != 0, cut[p], cut[p + 1]);
      (void)hipEventRecord(h->part_ev[p], st);
    }
    const bool packed = h->tune.int_layout >= 2 && h->pack_waves > 0;
    // first entry whose last bond is at or above bond c
    auto first_at = [](const std::vector<int64_t> &v, int64_t c) -> int {
      return (int)(std::lower_bound(v.begin(), v.end(), c) - v.begin());
    };
    // packed filaments on one side stream, the longer ones (their own
    // launches, latency-bound) on a second
    hipStream_t sl = h->side[1];
    for (int p = 0; p < P; ++p) {
      (void)hipStreamWaitEvent(si, h->part_ev[p], 0);
      (void)hipStreamWaitEvent(sl, h->part_ev[p], 0);
      const int64_t lo = cut[p], hi = p + 1 < P ? cut[p + 1] : nb + 1;
      if (packed)
        for (int c = 0; c < 3; ++c) {
          const int c0 = h->pack_cs[c], c1 = h->pack_cs[c + 1];
          auto at = [&](int64_t v) -> int {
            return (int)(std::lower_bound(h->pack_last.begin() + c0, h->pack_last.begin() + c1, v) -
                         h->pack_last.begin());
          };
          const int w0 = p == 0 ? c0 : at(lo);
          const int w1 = p + 1 < P ? at(hi) : c1;
          if (w1 > w0)
            sc::launch_integrate_packed(d, g, h->pack_tab + (int64_t)64 * w0, w1 - w0, 16 << c,
                                        h->tune.int_waves, odd, si);
        }
      for (const Bin &b : h->bins) {
        if (packed && b.maxn <= 128) continue;
        const int i0 = p == 0 ? 0 : first_at(b.last, lo);
        const int i1 = p + 1 < P ? first_at(b.last, hi) : b.count;
        if (i1 > i0)
          sc::launch_integrate(d, g, b.list + i0, i1 - i0, b.maxn, h->p.precision, h->tune,
                               odd, packed ? sl : si);
      }
    }
    for (int k = 0; k < 2; ++k) {
      (void)hipEventRecord(h->join_ev[k], h->side[k]);
      (void)hipStreamWaitEvent(st, h->join_ev[k], 0);
    }
    sc::launch_commit(d, g, st);
    return;
  }
  if (!odd) sc::launch_xlink_forces(d, g, st);
  sc::launch_pair(d, g, h->p.precision, h->tune, st, g.thermo != 0);
  if (!odd) {
    sc::launch_xlink_events(d, g, st);
    if (g.xpoff > 0) sc::launch_xlink_lists(d, g, st);
  }
  if (pf) {
    (void)hipEventRecord(h->ev[1], st);
    (void)hipEventRecord(h->ev[4], st);
  }
  // the integrator's launches: one per length bin, or, in the packed layout,
  // one for every filament of at most 128 sites (maxn 0 below) plus the
  // bin of longer ones
  sc::BinDesc bins[sc::kMaxBins + 1];
  int nbins = 0;
  const bool packed = h->tune.int_layout >= 2 && h->pack_waves > 0;
  // layout 2: one launch for every filament, the bin of 129-256 sites first
  const Bin *merged = nullptr;
  if (packed && h->tune.int_layout == 2) {
    for (const Bin &b : h->bins)
      if (b.maxn == 256) merged = &b;
    bins[nbins++] = {h->pack_tab, h->pack_waves, 0, 0};
  } else if (packed) {
    for (int c = 0; c < 3; ++c)
      if (h->pack_cs[c + 1] > h->pack_cs[c])
        bins[nbins++] = {h->pack_tab + (int64_t)64 * h->pack_cs[c],
                         h->pack_cs[c + 1] - h->pack_cs[c], 0, 16 << c};
  }
  for (const Bin &b : h->bins)
    if ((!packed || b.maxn > 128) && &b != merged) bins[nbins++] = {b.list, b.count, b.maxn, 0};
  auto load_of = [&](int k) -> int64_t {
    if (bins[k].maxn) return (int64_t)bins[k].count * bins[k].maxn;
    const int64_t l = (int64_t)bins[k].count * 120;
    return bins[k].width == 0 && merged ? l + (int64_t)merged->count * 256 : l;
  };
  auto launch_bin = [&](int k, hipStream_t sk) {
    if (bins[k].maxn == 0 && bins[k].width == 0 && merged)
      sc::launch_integrate_all(d, g, bins[k].list, bins[k].count, h->pack_cs[1], h->pack_cs[2],
                               merged->list, merged->count, h->tune.int_waves, odd, sk);
    else if (bins[k].maxn == 0 && bins[k].width == 0)
      sc::launch_integrate_packed_all(d, g, bins[k].list, bins[k].count, h->pack_cs[1],
                                      h->pack_cs[2], h->tune.int_waves, odd, sk);
    else if (bins[k].maxn == 0)
      sc::launch_integrate_packed(d, g, bins[k].list, bins[k].count, bins[k].width,
                                  h->tune.int_waves, odd, sk);
    else
      sc::launch_integrate(d, g, bins[k].list, bins[k].count, bins[k].maxn, h->p.precision,
                           h->tune, odd, sk);
  };
  if (h->tune.int_streams && nbins > 1 && h->side[0]) {
    // bins are independent (disjoint filaments; the displacement maxima
    // combine by max): deal them to the main and side streams, largest
    // first onto the least loaded, then join
    const int lanes = std::min(nbins, 1 + std::min(h->tune.int_streams, sc_system::kSide));
    int order[sc::kMaxBins + 1];
    for (int k = 0; k < nbins; ++k) order[k] = k;
    std::sort(order, order + nbins, [&](int a, int b) { return load_of(a) > load_of(b); });
    int64_t load[1 + sc_system::kSide] = {};
    int lane_of[sc::kMaxBins + 1];
    for (int j = 0; j < nbins; ++j) {
      int best = 0;
      for (int l = 1; l < lanes; ++l)
        if (load[l] < load[best]) best = l;
      lane_of[order[j]] = best;
      load[best] += load_of(order[j]);
    }
    (void)hipEventRecord(h->fork_ev, st);
    for (int l = 1; l < lanes; ++l) (void)hipStreamWaitEvent(h->side[l - 1], h->fork_ev, 0);
    for (int k = 0; k < nbins; ++k) {
      hipStream_t sk = lane_of[k] == 0 ? st : h->side[lane_of[k] - 1];
      launch_bin(k, sk);
    }
    for (int l = 1; l < lanes; ++l) {
      (void)hipEventRecord(h->join_ev[l - 1], h->side[l - 1]);
      (void)hipStreamWaitEvent(st, h->join_ev[l - 1], 0);
    }
  } else {
    for (int k = 0; k < nbins; ++k) launch_bin(k, st);
  }
  if (pf) {
    (void)hipEventRecord(h->ev[5], st);
    (void)hipEventRecord(h->ev[2], st);
  }
  sc::launch_commit(d, g, st);
  if (pf) (void)hipEventRecord(h->ev[3], st);
}

int capture_graphs(sc_system *h) {
  drop_graph(h);
  for (int k = 0; k < sc_system::kGraphs; ++k) {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int pairs = k == 3 ? sc_system::kGraphPairs : 1;
    for (int q = 0; q < pairs; ++q) {
      if (k == 0 || k >= 2) enqueue_step(h, true);
      if (k >= 1) enqueue_step(h, false);
    }
    hipGraph_t gr = nullptr;
    HIPCHK(h, hipStreamEndCapture(h->stream, &gr));
    h->graph[k] = gr;
    HIPCHK(h, hipGraphInstantiate(&h->gexec[k], gr, nullptr, nullptr, 0));
  }
  return SC_OK;
}

// Runs the rebuild stage outside a step (load time), growing the candidate
// capacity until the list fits.
int initial_rebuild(sc_system *h) {
  for (int attempt = 0; attempt < 16; ++attempt) {
    Ctl c = read_ctl(h);
    c.rebuild = 1;
    c.abort = 0;
    c.overflow = 0;
    int rc = write_ctl(h, c);
    if (rc) return rc;
    sc::launch_rebuild(h->d, h->g, h->stream);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipGetLastError());
    c = read_ctl(h);
    if (c.overflow) {
      if ((rc = alloc_cand(h, grown_capacity(h->g.cap, c.nc_need)))) return rc;
      continue;
    }
    c.rebuild = 0;
    return write_ctl(h, c);
  }
  return fail(h, SC_ERR_NOMEM, "candidate capacity growth did not converge");
}

int check_handle(sc_system *h, bool need_loaded) {
  if (!h) return SC_ERR_ARG;
  if (need_loaded && !h->loaded)
    return fail(h, SC_ERR_STATE, "no filaments loaded");
  return SC_OK;
}

void get3(double *dst, const double *src_dev, int64_t n) {
  (void)hipMemcpy(dst, src_dev, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
}

// storage-order array of `comp` components x n entries -> caller order
template <typename T>
void to_caller(const sc_system *h, const std::vector<int64_t> &map, const T *src, T *dst,
               int64_t n, int comp) {
  if (!h->reordered) {
    std::copy(src, src + comp * n, dst);
    return;
  }
  for (int k = 0; k < comp; ++k)
    for (int64_t i = 0; i < n; ++i) dst[k * n + map[i]] = src[k * n + i];
}

// device SoA site/bond array -> caller-order host array
// Blocks of the pair pass at its smallest block size (64 bonds): each owns
// one stress slot, whatever launch shape is set while the sum accumulates.
int64_t stress_slots(int64_t nb) { return nb / 64 + 2; }

int fetch_mapped(sc_system *h, const double *dev, double *out, int64_t n, int comp,
                 bool sites) {
  std::vector<double> tmp((size_t)comp * n);
  HIPCHK(h, hipMemcpy(tmp.data(), dev, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
  to_caller(h, sites ? h->site_map : h->bond_map, tmp.data(), out, n, comp);
  return SC_OK;
}

// Space-filling-curve order of the filaments' centres: 10-bit Morton code
// of the wrapped scaled centre per dimension, stable in caller order.  The
// top 3 bits per dimension (512 blocks) come first, then the integrator's
// length bin, then the rest of the code: inside a block each bin's filaments
// are contiguous, so one bin's launch does not fetch cache lines that
// mostly hold another bin's sites, while a bond's partners stay in its block.
std::vector<int32_t> spatial_permutation(const sc_params &p, int32_t nf, const int32_t *so,
                                         const double *r, std::vector<int32_t> *block) {
  const int64_t ns = so[nf];
  std::vector<uint64_t> key(nf);
  for (int f = 0; f < nf; ++f) {
    double c[3] = {0, 0, 0};
    for (int64_t i = so[f]; i < so[f + 1]; ++i)
      for (int k = 0; k < 3; ++k) c[k] += r[k * ns + i];
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
      double sk = 0;
      for (int j = 0; j < 3; ++j) sk += p.unit_cell_inv[3 * k + j] * c[j] / (so[f + 1] - so[f]);
      sk -= std::floor(sk + 0.5);
      uint32_t q = (uint32_t)std::min(1023.0, std::max(0.0, std::floor((sk + 0.5) * 1024.0)));
      for (int b = 0; b < 10; ++b) code |= ((q >> b) & 1u) << (3 * b + k);
    }
    const int n = so[f + 1] - so[f];
    uint64_t bin = 0;
    while (bin < 5 && n > (8 << bin)) ++bin;  // bins of 8, 16, ..., 256 sites
    // filaments of more than 128 sites first: their bonds then fall in the
    // first bond range, so with the pair pass and the integrator
    // overlapped, their (latency-bound) integrator starts earliest
    const uint64_t late = n > 128 ? 0 : 1;
    key[f] = (late << 42) | ((uint64_t)(code >> 21) << 33) | (bin << 30) | code;
  }
  std::vector<int32_t> perm(nf);
  for (int f = 0; f < nf; ++f) perm[f] = f;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  if (block) {
    block->resize(nf);
    for (int j = 0; j < nf; ++j) (*block)[j] = (int32_t)(key[perm[j]] >> 33);
  }
  return perm;
}

// Packed integrator layout (k_integrate_packed): each filament of at most
// 128 sites takes w = ceil(n/2) consecutive lanes of a wavefront.  A segment
// of 33-64 lanes starts a wave at lane 0; one of 17-32 lanes starts at lane
// 0 or 32 with both of that half's 16-lane rows otherwise empty; one of at
// most 16 lanes lies inside a single row (best fit over the rows' free
// tails, widest filaments first).  The lane moves of a wave are those of
// its widest class (16-, 32- or 64-lane groups), valid for every segment
// in it because no segment crosses a group boundary of its class.
struct PackWave {
  int code = 0;                   // 0: 16-lane rows, 1: 32-lane halves, 2: 64
  int32_t lane[64];               // filament (caller-side id) per lane, -1 idle
  int gl[64];                     // lane inside the segment
  int row_used[4] = {0, 0, 0, 0};  // lanes used from the start of each row
  PackWave() {
    for (int l = 0; l < 64; ++l) {
      lane[l] = -1;
      gl[l] = 0;
    }
  }
};

void pack_place(PackWave &w, int32_t id, int start, int width) {
  for (int k = 0; k < width; ++k) {
    w.lane[start + k] = id;
    w.gl[start + k] = k;
  }
  for (int r = 0; r < 4; ++r) {
    const int lo = std::max(start, 16 * r), hi = std::min(start + width, 16 * r + 16);
    if (hi > lo) w.row_used[r] = std::max(w.row_used[r], hi - 16 * r);
  }
}

// Packs `ids` (width = lanes each, 1..64) into waves appended to `out`.
void pack_group(const std::vector<int32_t> &ids, const std::vector<int> &width,
                std::vector<PackWave> &out) {
  std::vector<int32_t> order(ids);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return width[a] > width[b]; });
  const size_t first = out.size();
  // rows with free lanes, by free count (best fit for the narrow segments)
  std::vector<std::vector<std::pair<size_t, int>>> rows_by_free(17);
  std::vector<std::pair<size_t, int>> free_halves;  // (wave, half)
  auto add_rows = [&](size_t wi) {
    const PackWave &w = out[wi];
    for (int r = 0; r < 4; ++r) {
      const int fr = 16 - w.row_used[r];
      if (fr > 0 && fr < 16) rows_by_free[fr].push_back({wi, r});
    }
  };
  for (int32_t id : order) {
    const int wd = width[id];
    if (wd > 32) {
      out.emplace_back();
      out.back().code = 2;
      pack_place(out.back(), id, 0, wd);
      const size_t wi = out.size() - 1;
      add_rows(wi);
      for (int r = (wd + 15) / 16; r < 4; ++r) rows_by_free[16].push_back({wi, r});
    } else if (wd > 16) {
      size_t wi;
      int half;
      if (!free_halves.empty()) {
        wi = free_halves.back().first;
        half = free_halves.back().second;
        free_halves.pop_back();
      } else {
        out.emplace_back();
        wi = out.size() - 1;
        half = 0;
        free_halves.push_back({wi, 1});
      }
      PackWave &w = out[wi];
      w.code = std::max(w.code, 1);
      pack_place(w, id, 32 * half, wd);
      const int r = 2 * half + 1;  // the half's second row, partly used
      const int fr = 16 - w.row_used[r];
      if (fr > 0) rows_by_free[fr].push_back({wi, r});
    } else {
      // best fit: the row with the fewest free lanes that still holds wd
      size_t wi = 0;
      int r = -1;
      for (int fr = wd; fr <= 16 && r < 0; ++fr) {
        while (!rows_by_free[fr].empty()) {
          auto cand = rows_by_free[fr].back();
          rows_by_free[fr].pop_back();
          if (16 - out[cand.first].row_used[cand.second] != fr) continue;  // stale
          wi = cand.first;
          r = cand.second;
          break;
        }
      }
      if (r < 0 && !free_halves.empty()) {
        // an untouched half of a 32-lane wave
        wi = free_halves.back().first;
        const int half = free_halves.back().second;
        free_halves.pop_back();
        r = 2 * half;
        rows_by_free[16].push_back({wi, 2 * half + 1});
      }
      if (r < 0) {
        out.emplace_back();
        wi = out.size() - 1;
        r = 0;
        for (int q = 1; q < 4; ++q) rows_by_free[16].push_back({wi, q});
      }
      PackWave &w = out[wi];
      pack_place(w, id, 16 * r + w.row_used[r], wd);
      const int fr = 16 - w.row_used[r];
      if (fr > 0) rows_by_free[fr].push_back({wi, r});
    }
  }
  (void)first;
}

// Enqueues `count` steps after step index `from` (the next step is
// from+1), then the Ctl mirror copy and the completion event.
int enqueue_range(sc_system *h, int64_t from, int64_t count) {
  const bool use_graph = h->p.use_graph && !h->prof;
  int rc;
  if (use_graph && !h->gexec[0])
    if ((rc = capture_graphs(h))) return rc;
  int64_t next = from + 1;
  int64_t s = 0;
  while (s < count) {
    const bool odd = (next & 1) != 0;
    if (use_graph) {
      if (odd && count - s >= 2 * sc_system::kGraphPairs) {
        HIPCHK(h, hipGraphLaunch(h->gexec[3], h->stream));
        s += 2 * sc_system::kGraphPairs;
        next += 2 * sc_system::kGraphPairs;
        continue;
      }
      if (odd && count - s >= 2) {
        HIPCHK(h, hipGraphLaunch(h->gexec[2], h->stream));
        s += 2;
        next += 2;
        continue;
      }
      HIPCHK(h, hipGraphLaunch(h->gexec[odd ? 0 : 1], h->stream));
    } else {
      enqueue_step(h, odd);
      if (h->prof) {
        HIPCHK(h, hipEventSynchronize(h->ev[3]));
        float ms = 0.f;
        const int pairs[kFamilies][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}};
        for (int k = 0; k < kFamilies; ++k) {
          if (k == 3 && !odd) continue;
          (void)hipEventElapsedTime(&ms, h->ev[pairs[k][0]], h->ev[pairs[k][1]]);
          h->prof_ms[k] += ms;
          h->prof_n[k] += 1;
        }
      }
    }
    s += 1;
    next += 1;
  }
  HIPCHK(h, hipMemcpyAsync(h->ctl_host, h->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipEventRecord(h->done_ev, h->stream));
  h->pending = true;
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_abi_version(void) { return SC_ABI_VERSION; }

int sc_create(const sc_params *params, sc_handle *out) {
  return sc::system_create(params, nullptr, nullptr, out);
}

}  // extern "C"

namespace sc {

int system_create(const sc_params *params, hipStream_t stream, Arena *arena,
                  sc_handle *out, Ctl *pinned_ctl) {
  if (!params || !out) return SC_ERR_ARG;
  *out = nullptr;
  const sc_params &p = *params;
  if (p.n_dim != 3 || p.n_periodic != 3) return SC_ERR_ARG;
  if (!(p.delta > 0) || !(p.system_radius > 0) || !(p.wca_sig > 0))
    return SC_ERR_ARG;
  if (p.precision != SC_PREC_FP64 && p.precision != SC_PREC_MIXED)
    return SC_ERR_ARG;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (i != j && (p.unit_cell[3 * i + j] != 0.0 || p.unit_cell_inv[3 * i + j] != 0.0))
        return SC_ERR_ARG;  // orthorhombic boxes only
  {
    const int64_t n = cells_per_side(p);
    if (n * n * n >= (int64_t)1 << 31) return SC_ERR_ARG;
    // the mixed path's compact bond records pack 10 bits per cell coordinate
    if (p.precision == SC_PREC_MIXED && n > 1024) return SC_ERR_ARG;
  }
  int ndev = 0, device = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SC_ERR_HIP;
  if (hipGetDevice(&device) != hipSuccess) return SC_ERR_HIP;
  // the allocation-free histogram's scratch address on this device
  if (scratch_setup(device) != hipSuccess) return SC_ERR_HIP;
  sc_system *h = new sc_system();
  h->device = device;
  h->p = p;
  Geo &g = h->g;
  std::memcpy(g.h, p.unit_cell, sizeof(g.h));
  std::memcpy(g.hinv, p.unit_cell_inv, sizeof(g.hinv));
  // WCAPotential::Init (include/simcore/wca_potential.hpp:34-48)
  double rcut = std::pow(2.0, 1.0 / 6.0) * p.wca_sig;
  g.rcut2 = rcut * rcut;
  g.c12 = 4.0 * p.wca_eps * std::pow(p.wca_sig, 12.0);
  g.c6 = 4.0 * p.wca_eps * std::pow(p.wca_sig, 6.0);
  g.eps = p.wca_eps;
  g.fcut = p.f_cutoff;
  g.delta = p.delta;
  const int n1d = cells_per_side(p);
  double cell_len = (double)2 * p.system_radius / n1d;
  g.n1d = n1d;
  g.thr = 0.25 * cell_len * cell_len;
  g.n_cells = (int64_t)n1d * n1d * n1d;
  g.occ_words = (n1d + 31) / 32;
  build_home_order(g);
  g.stoch = p.stoch_flag ? 1 : 0;
  g.seed = p.seed;
  // mixed precision keeps 32-byte bond records whose cell coordinates are
  // packed in 10 bits each (at most 1024 cells per side, checked above)
  g.compact = p.precision == SC_PREC_MIXED ? 1 : 0;
  g.inv_n1d = 1.0 / n1d;
  h->arena = arena;
  if (stream) {
    h->stream = stream;
    h->own_stream = false;
  } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return SC_ERR_HIP;
  }
  void *pin = pinned_ctl;
  h->own_ctl_host = pinned_ctl == nullptr;
  if (hipEventCreateWithFlags(&h->done_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming) != hipSuccess ||
      (!pin && hipHostMalloc(&pin, sizeof(Ctl), hipHostMallocDefault) != hipSuccess)) {
    sc_destroy(h);
    return SC_ERR_HIP;
  }
  h->ctl_host = static_cast<Ctl *>(pin);
  *h->ctl_host = Ctl{};
  *out = h;
  return SC_OK;
}

bool system_batchable(sc_handle h) {
  if (!h || !h->loaded || h->pending || h->prof || h->p.precision != SC_PREC_MIXED ||
      h->g.nx != 0 || h->g.thermo || h->g.xi_buf || !h->rec[0] || h->ctl_host->abort ||
      (h->i_step_host & 1) != 0)
    return false;
  // every filament in the packed waves or the 256-site bin
  for (const Bin &b : h->bins)
    if (b.maxn > 128 && b.maxn != 256) return false;
  return h->pack_waves > 0 || !h->bins.empty();
}

void system_member(sc_handle h, Member *m) {
  m->d[0] = h->d;
  m->d[0].br = h->rec[0];
  m->d[0].br_out = h->rec[1];
  m->d[1] = h->d;
  m->d[1].br = h->rec[1];
  m->d[1].br_out = h->rec[0];
  m->g = h->g;
  m->pack_tab = h->pack_tab;
  m->pack_waves = h->pack_waves;
  m->cs1 = h->pack_cs[1];
  m->cs2 = h->pack_cs[2];
  m->long_list = nullptr;
  m->n_long = 0;
  for (const Bin &b : h->bins)
    if (b.maxn == 256) {
      m->long_list = b.list;
      m->n_long = b.count;
    }
}

int system_batch_done(sc_handle h, const Ctl &c, int64_t target, int *done) {
  // the batch's range becomes this handle's pending range, completed on
  // the (idle) system stream, and is polled as such
  *h->ctl_host = c;
  h->target = target;
  h->grow_rounds = 0;
  h->pending = true;
  HIPCHK(h, hipEventRecord(h->done_ev, h->stream));
  return sc_step_poll(h, done);
}

size_t system_estimate_bytes(const sc_params *params, int32_t nf, int64_t ns, int n_bins) {
  const int64_t n1d = cells_per_side(*params);
  const int64_t n_cells = n1d * n1d * n1d;
  const int64_t nb = ns - nf;
  auto R = [](int64_t n) { return Arena::round((size_t)(n > 0 ? n : 1)); };
  size_t b = R(sizeof(Ctl)) + R(8 * sc::kDrSlots) + R(4 * (nf + 1)) + R(4 * (int64_t)nf) +
             R(64 * (int64_t)nf) +
             R(8 * (int64_t)nf) + R(4 * nb);
  b += 4 * R(24 * ns) + 2 * R(8 * ns);                    // r, rp, fr, dr0, acc, acc_prev
  const int64_t rec_bytes = params->precision == SC_PREC_MIXED ? 32 : 64;
  b += R(rec_bytes * nb) + 2 * R(24 * nb) + R(8 * nb);    // record, F, T (at most), tension
  const int64_t ncol = n1d * n1d, occ_words = (n1d + 31) / 32;
  b += R(4 * nb) + R(4 * ncol) + R(4 * (ncol + 1)) + 2 * R(4 * nb);  // keys, columns, col_tmp/rank
  b += R(8 * n_cells) + R(4 * ncol * occ_words);                 // cell runs, occupancy
  b += R(2 * ncol * occ_words) + R(4 * ncol) + R(4 * nb);        // run tables
  b += 2 * R(4 * nb) + R(4 * nb) + R(4 * (nb + 1));  // sorted, skey, deg, inc_start
  b += R(8 * sc::kRunSlab * nb);                      // run slabs
  const int64_t tmp = std::max(scan_tmp_elems(ncol), scan_tmp_elems(nb));
  b += R(4 * tmp);
  b += R(4 * (int64_t)nf) + (size_t)n_bins * Arena::kAlign;  // bin lists
  if (params->precision == SC_PREC_MIXED)  // packed layout: at most one wave per filament
    b += R(4 * 64 * (int64_t)nf) + R(4 * (int64_t)nf);
  const int64_t cap = initial_capacity(*params, nb);
  const int64_t slot = params->precision == SC_PREC_FP64 ? 32 : 16;
  b += R(16 * cap) + R(2 * slot * cap);
  return b;
}

}  // namespace sc

namespace {

struct HostCand {
  int32_t key, phase, rank1, rank2, b1, b2;
  int64_t q;  // the role-0 incidence (the candidate's slot copy)
};

// The candidate list in the reference's order, recovered from the
// incidence lists: each role-0 incidence of bond b with partner p is the
// candidate (b, p), and CellList::MakePairs (src/cell_list.cpp:127-137,
// src/cell.cpp:14-70) orders candidates by the cell key of obj1, then phase
// (the cell's self pairs, then its 13 half-shell cells in stencil order),
// then obj1's rank in its cell, then obj2's rank in its cell, ranks in
// storage order.  Storage indices.
int host_candidates(sc_system *h, std::vector<HostCand> *out, std::vector<int2> *inc_out) {
  const int64_t nb = h->g.nb;
  const int n = h->g.n1d;
  const int64_t ni = 2 * read_ctl(h).nc;
  std::vector<int32_t> keys(nb), start(nb + 1);
  std::vector<int2> inc(ni);
  HIPCHK(h, hipMemcpy(keys.data(), h->d.keys, sizeof(int32_t) * nb, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(start.data(), h->d.inc_start, sizeof(int32_t) * (nb + 1),
                      hipMemcpyDeviceToHost));
  if (ni) HIPCHK(h, hipMemcpy(inc.data(), h->d.inc, sizeof(int2) * ni, hipMemcpyDeviceToHost));
  if (start[nb] != ni) return fail(h, SC_ERR_STATE, "incidence lists and list length disagree");
  std::vector<int32_t> order(nb), rank(nb);
  for (int64_t b = 0; b < nb; ++b) order[b] = (int32_t)b;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
  for (int64_t j = 0; j < nb; ++j)
    rank[order[j]] = (j > 0 && keys[order[j - 1]] == keys[order[j]]) ? rank[order[j - 1]] + 1 : 0;
  out->clear();
  out->reserve(ni / 2);
  for (int64_t b = 0; b < nb; ++b) {
    int32_t sh[13];
    for (int q = 0; q < 13; ++q) sh[q] = host_shell_key(n, keys[b], q);
    for (int64_t q = start[b]; q < start[b + 1]; ++q) {
      if (inc[q].y & 1) continue;
      const int32_t p = inc[q].x;
      int phase = -1;
      if (keys[p] == keys[b]) {
        phase = 0;
      } else {
        for (int k = 0; k < 13; ++k)
          if (sh[k] == keys[p]) phase = 1 + k;
      }
      if (phase < 0) return fail(h, SC_ERR_STATE, "incidence outside the half shell");
      out->push_back({keys[b], phase, rank[b], rank[p], (int32_t)b, p, q});
    }
  }
  std::sort(out->begin(), out->end(), [](const HostCand &a, const HostCand &b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.phase != b.phase) return a.phase < b.phase;
    if (a.rank1 != b.rank1) return a.rank1 < b.rank1;
    return a.rank2 < b.rank2;
  });
  if (inc_out) inc_out->swap(inc);
  return SC_OK;
}

}  // namespace

extern "C" {

int sc_load_filaments(sc_handle h, int32_t nf, const int32_t *site_off,
                      const double *r, const sc_filament_params *fil) {
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (h->loaded) return fail(h, SC_ERR_STATE, "filaments already loaded");
  if (nf <= 0 || !site_off || !r || !fil) return fail(h, SC_ERR_ARG, "null input");
  if (site_off[0] != 0) return fail(h, SC_ERR_ARG, "site_off[0] must be 0");
  for (int f = 0; f < nf; ++f) {
    int n = site_off[f + 1] - site_off[f];
    if (n < 3) return fail(h, SC_ERR_ARG, "a filament needs at least 2 bonds");
    if (n > 256) return fail(h, SC_ERR_ARG, "at most 256 sites per filament");
    if (!(fil[f].bond_length > 0) || !(fil[f].gamma_par > 0) || !(fil[f].gamma_perp > 0))
      return fail(h, SC_ERR_ARG, "bad filament parameters");
  }
  Geo &g = h->g;
  Dev &d = h->d;
  g.nf = nf;
  g.ns = site_off[nf];
  g.nb = g.ns - nf;
  // storage order (caller order unless spatial_order is set)
  std::vector<int32_t> so_st, noise_off(nf);
  std::vector<double> r_st;
  std::vector<sc_filament_params> fil_st;
  // packed integrator layout (mixed precision, filaments of <= 128 sites):
  // groups of filaments packed together -- the Morton blocks in spatial
  // order, everything in caller order -- and, in spatial order, the
  // storage order follows the packing so each wave reads contiguous sites
  const bool packing = h->p.precision == SC_PREC_MIXED && nf < (1 << 23);
  std::vector<PackWave> waves;
  std::vector<int> width(packing ? nf : 0);
  for (int f = 0; packing && f < nf; ++f) {
    const int n = site_off[f + 1] - site_off[f];
    width[f] = n <= 128 ? (n + 1) / 2 : 0;
  }
  if (h->p.spatial_order) {
    const int64_t ns_ = g.ns, nb_ = g.nb;
    std::vector<int32_t> block;
    h->perm = spatial_permutation(h->p, nf, site_off, r, &block);
    if (packing) {
      std::vector<int32_t> order;
      order.reserve(nf);
      for (int j0 = 0; j0 < nf;) {
        int j1 = j0;
        while (j1 < nf && block[j1] == block[j0]) ++j1;
        std::vector<int32_t> ids, longs;
        for (int j = j0; j < j1; ++j)
          (width[h->perm[j]] ? ids : longs).push_back(h->perm[j]);
        const size_t w0 = waves.size();
        pack_group(ids, width, waves);
        for (size_t wi = w0; wi < waves.size(); ++wi)
          for (int l = 0; l < 64; ++l)
            if (waves[wi].lane[l] >= 0 && waves[wi].gl[l] == 0) order.push_back(waves[wi].lane[l]);
        order.insert(order.end(), longs.begin(), longs.end());
        j0 = j1;
      }
      h->perm = order;
    }
    h->reordered = true;
    so_st.assign(nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
      const int u = h->perm[f];
      so_st[f + 1] = so_st[f] + (site_off[u + 1] - site_off[u]);
    }
    r_st.resize(3 * (size_t)ns_);
    fil_st.resize(nf);
    h->site_map.resize(ns_);
    h->bond_map.resize(nb_);
    for (int f = 0; f < nf; ++f) {
      const int u = h->perm[f];
      fil_st[f] = fil[u];
      noise_off[f] = site_off[u];
      const int n = so_st[f + 1] - so_st[f];
      for (int i = 0; i < n; ++i) {
        h->site_map[so_st[f] + i] = site_off[u] + i;
        for (int k = 0; k < 3; ++k) r_st[k * ns_ + so_st[f] + i] = r[k * ns_ + site_off[u] + i];
        if (i < n - 1) h->bond_map[so_st[f] - f + i] = site_off[u] - u + i;
      }
    }
    site_off = so_st.data();
    r = r_st.data();
    fil = fil_st.data();
  } else {
    for (int f = 0; f < nf; ++f) noise_off[f] = site_off[f];
    if (packing) {
      std::vector<int32_t> ids;
      for (int f = 0; f < nf; ++f)
        if (width[f]) ids.push_back(f);
      pack_group(ids, width, waves);
    }
  }
  h->site_off.assign(site_off, site_off + nf + 1);
  const int64_t ns = g.ns, nb = g.nb;
  if ((rc = dalloc(h, &d.ctl, 1, false))) return rc;
  if ((rc = dalloc(h, &d.dr_slots, sc::kDrSlots, false))) return rc;
  HIPCHK(h, hipMemsetAsync(d.dr_slots, 0, sizeof(unsigned long long) * sc::kDrSlots,
                           h->stream));
  if ((rc = dalloc(h, &d.site_off, nf + 1, false))) return rc;
  if ((rc = dalloc(h, &d.noise_off, nf, false))) return rc;
  if ((rc = dalloc(h, &d.fpar, 8 * (size_t)nf, false))) return rc;
  if ((rc = dalloc(h, &d.norm, 2 * (size_t)nf, false))) return rc;
  if ((rc = dalloc(h, &d.bond_fil, nb, false))) return rc;
  double **site_arrays[] = {&d.r, &d.rp, &d.fr, &d.dr0};
  for (double **a : site_arrays)
    if ((rc = dalloc(h, a, 3 * (size_t)ns, false))) return rc;
  if ((rc = dalloc(h, &d.acc, ns, false))) return rc;
  if ((rc = dalloc(h, &d.acc_prev, ns, false))) return rc;
  for (int k = 0; k < 2; ++k)
    if ((rc = dalloc(h, &h->rec[k], (g.compact ? 4 : 8) * (size_t)nb, false))) return rc;
  d.br = d.br_out = h->rec[0];
  if (h->p.precision == SC_PREC_FP64) {
    if ((rc = dalloc(h, &d.bf, 3 * (size_t)nb, false))) return rc;
    if ((rc = dalloc(h, &d.bt, 3 * (size_t)nb, false))) return rc;
  } else {
    if ((rc = dalloc(h, &d.bf32, 3 * (size_t)nb, false))) return rc;
    if ((rc = dalloc(h, &d.bt32, 3 * (size_t)nb, false))) return rc;
  }
  if ((rc = dalloc(h, &d.tension, nb, false))) return rc;
  if ((rc = dalloc(h, &d.bpote, nb, false))) return rc;
  if ((rc = dalloc(h, &d.stress_acc, 6 * (size_t)stress_slots(nb), false))) return rc;
  if ((rc = dalloc(h, &d.keys, nb, false))) return rc;
  const int64_t ncol = (int64_t)g.n1d * g.n1d;
  if ((rc = dalloc(h, &d.col_count, ncol, false))) return rc;
  if ((rc = dalloc(h, &d.col_start, ncol + 1, false))) return rc;
  if ((rc = dalloc(h, &d.col_tmp, nb, false))) return rc;
  if ((rc = dalloc(h, &d.col_rank, nb, false))) return rc;
  if ((rc = dalloc(h, &d.cell_run, g.n_cells, false))) return rc;
  if ((rc = dalloc(h, &d.sorted, nb, false))) return rc;
  if ((rc = dalloc(h, &d.skey, nb, false))) return rc;
  if ((rc = dalloc(h, &d.cell_occ, ncol * g.occ_words, false))) return rc;
  if ((rc = dalloc(h, &d.occ_pre, ncol * g.occ_words, false))) return rc;
  if ((rc = dalloc(h, &d.col_ncell, ncol, false))) return rc;
  if ((rc = dalloc(h, &d.cell_first, nb, false))) return rc;
  if ((rc = dalloc(h, &d.deg, nb, false))) return rc;
  if ((rc = dalloc(h, &d.nrun, sc::kRunSlab * (size_t)nb, false))) return rc;
  if ((rc = dalloc(h, &d.inc_start, nb + 1, false))) return rc;
  int64_t tmp = std::max(sc::scan_tmp_elems(ncol), sc::scan_tmp_elems(nb));
  if ((rc = dalloc(h, &d.scan_tmp, tmp, false))) return rc;

  // host-side per-filament tables
  std::vector<double> fpar(8 * (size_t)nf);
  std::vector<int32_t> norm(2 * (size_t)nf), bond_fil(nb);
  int max_n = 0;
  for (int f = 0; f < nf; ++f) {
    const sc_filament_params &q = fil[f];
    double *fp = &fpar[8 * (size_t)f];
    fp[0] = q.bond_length;
    h->max_bond_length = std::max(h->max_bond_length, q.bond_length);
    fp[1] = q.persistence_length;
    fp[2] = q.gamma_par;
    fp[3] = q.gamma_perp;
    // Filament::SetDiffusion (src/filament.cpp:346-347)
    fp[4] = std::sqrt(24.0 * q.gamma_par / h->p.delta);
    fp[5] = std::sqrt(24.0 * q.gamma_perp / h->p.delta);
    fp[6] = q.diameter;
    fp[7] = 0.0;
    norm[2 * f] = 1;  // normalize_switch_ starts true
    norm[2 * f + 1] = 0;
    int n = site_off[f + 1] - site_off[f];
    max_n = std::max(max_n, n);
    for (int i = 0; i < n - 1; ++i) bond_fil[site_off[f] - f + i] = f;
  }
  HIPCHK(h, hipMemcpyAsync(d.site_off, site_off, sizeof(int32_t) * (nf + 1), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.noise_off, noise_off.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.fpar, fpar.data(), sizeof(double) * fpar.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.norm, norm.data(), sizeof(int32_t) * norm.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.bond_fil, bond_fil.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.r, r, sizeof(double) * 3 * ns, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.rp, r, sizeof(double) * 3 * ns, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(d.dr0, r, sizeof(double) * 3 * ns, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(d.fr, 0, sizeof(double) * 3 * ns, h->stream));
  HIPCHK(h, hipMemsetAsync(d.acc, 0, sizeof(double) * ns, h->stream));
  HIPCHK(h, hipMemsetAsync(d.acc_prev, 0, sizeof(double) * ns, h->stream));
  if (d.bf) {
    HIPCHK(h, hipMemsetAsync(d.bf, 0, sizeof(double) * 3 * nb, h->stream));
    HIPCHK(h, hipMemsetAsync(d.bt, 0, sizeof(double) * 3 * nb, h->stream));
  } else {
    HIPCHK(h, hipMemsetAsync(d.bf32, 0, sizeof(float) * 3 * nb, h->stream));
    HIPCHK(h, hipMemsetAsync(d.bt32, 0, sizeof(float) * 3 * nb, h->stream));
  }
  HIPCHK(h, hipMemsetAsync(d.tension, 0, sizeof(double) * nb, h->stream));
  HIPCHK(h, hipMemsetAsync(d.bpote, 0, sizeof(double) * nb, h->stream));
  HIPCHK(h, hipMemsetAsync(d.stress_acc, 0, sizeof(double) * 6 * stress_slots(nb), h->stream));
  HIPCHK(h, hipMemsetAsync(d.col_count, 0, sizeof(int32_t) * ncol, h->stream));
  HIPCHK(h, hipMemsetAsync(d.cell_occ, 0, sizeof(uint32_t) * ncol * g.occ_words, h->stream));
  HIPCHK(h, hipMemsetAsync(d.deg, 0, sizeof(int32_t) * nb, h->stream));
  Ctl c{};
  if ((rc = write_ctl(h, c))) return rc;

  // length bins for the integrator (one launch each)
  const int bin_max[] = {8, 16, 32, 64, 128, 256};
  constexpr int kBins = 6;
  std::vector<std::vector<int32_t>> lists(kBins);
  for (int f = 0; f < nf; ++f) {
    int n = site_off[f + 1] - site_off[f];
    for (int k = 0; k < kBins; ++k)
      if (n <= bin_max[k]) {
        lists[k].push_back(f);
        break;
      }
  }
  for (int k = 0; k < kBins; ++k) {
    if (lists[k].empty()) continue;
    Bin b;
    b.maxn = bin_max[k];
    b.count = (int)lists[k].size();
    for (int32_t f : lists[k]) b.last.push_back((int64_t)site_off[f + 1] - f - 2);
    if (b.maxn > 128) h->long_last = std::max(h->long_last, b.last.back());
    if ((rc = dalloc(h, &b.list, lists[k].size(), false))) return rc;
    HIPCHK(h, hipMemcpyAsync(b.list, lists[k].data(), sizeof(int32_t) * lists[k].size(),
                             hipMemcpyHostToDevice, h->stream));
    h->bins.push_back(b);
  }
  // (host arrays of the packed layout live until the copies below have run)
  std::vector<int32_t> tab;
  if (packing && !waves.empty()) {
    // caller-side ids -> storage indices
    std::vector<int32_t> st_of(nf);
    for (int f = 0; f < nf; ++f) st_of[h->reordered ? h->perm[f] : f] = f;
    // waves grouped by lane-group width (one launch each), inside a group
    // in the order of their filaments' last bond, so that the waves whose
    // bonds all lie below a cut form a prefix of the group (overlapped
    // steps)
    {
      std::vector<int64_t> wl(waves.size(), -1);
      for (size_t wi = 0; wi < waves.size(); ++wi)
        for (int l = 0; l < 64; ++l)
          if (waves[wi].lane[l] >= 0) {
            const int32_t f = st_of[waves[wi].lane[l]];
            wl[wi] = std::max<int64_t>(wl[wi], (int64_t)site_off[f + 1] - f - 2);
          }
      std::vector<size_t> ord(waves.size());
      for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
      std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return waves[a].code != waves[b].code ? waves[a].code < waves[b].code : wl[a] < wl[b];
      });
      std::vector<PackWave> sorted_waves;
      sorted_waves.reserve(waves.size());
      h->pack_last.clear();
      for (int c = 0; c < 4; ++c) h->pack_cs[c] = 0;
      for (size_t i : ord) {
        sorted_waves.push_back(waves[i]);
        h->pack_last.push_back(wl[i]);
        for (int c = waves[i].code + 1; c < 4; ++c) h->pack_cs[c] += 1;
      }
      waves.swap(sorted_waves);
    }
    const size_t nw = waves.size();
    tab.assign(64 * nw, -1);
    for (size_t wi = 0; wi < nw; ++wi) {
      for (int l = 0; l < 64; ++l) {
        const int32_t id = waves[wi].lane[l];
        tab[64 * wi + l] = id < 0 ? -1 : ((st_of[id] << 8) | (waves[wi].code << 6) | waves[wi].gl[l]);
        if (id >= 0 && waves[wi].gl[l] == 0) h->pack_sites += site_off[st_of[id] + 1] - site_off[st_of[id]];
      }
    }
    if ((rc = dalloc(h, &h->pack_tab, tab.size(), false))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->pack_tab, tab.data(), sizeof(int32_t) * tab.size(),
                             hipMemcpyHostToDevice, h->stream));
    h->pack_waves = (int)nw;
    h->tune.int_layout = 2;
  }
  if ((rc = alloc_cand(h, initial_capacity(h->p, nb)))) return rc;
  sc::launch_bond_init(d, g, h->p.precision, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  if ((rc = initial_rebuild(h))) return rc;
  sc::launch_decide(d, g, h->stream);  // the decision for step 1
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *h->ctl_host = read_ctl(h);
  h->i_step_host = h->ctl_host->i_step;
  h->loaded = true;
  return SC_OK;
}

int sc_load_crosslinks(sc_handle h, int64_t n, const int32_t *bonds, const double *lambda,
                       const sc_crosslink_params *p) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!p || n < 0 || (n > 0 && (!bonds || !lambda))) return fail(h, SC_ERR_ARG, "null input");
  if (h->g.nx > 0) return fail(h, SC_ERR_STATE, "crosslinks already loaded");
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  if (n == 0) return SC_OK;
  const int64_t nb = h->g.nb;
  if (n > ((int64_t)1 << 30)) return fail(h, SC_ERR_ARG, "too many crosslinks");
  // caller bond numbering -> storage
  std::vector<int32_t> to_storage(nb), caller(nb);
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t c = h->reordered ? h->bond_map[b] : b;
    caller[b] = (int32_t)c;
    to_storage[c] = (int32_t)b;
  }
  std::vector<int32_t> xb(2 * n);
  for (int64_t i = 0; i < 2 * n; ++i) {
    if (bonds[i] < 0 || bonds[i] >= nb) return fail(h, SC_ERR_ARG, "crosslink bond out of range");
    if (!(lambda[i] >= 0)) return fail(h, SC_ERR_ARG, "negative anchor position");
    xb[i] = to_storage[bonds[i]];
  }
  for (int64_t x = 0; x < n; ++x)
    if (xb[2 * x] == xb[2 * x + 1]) return fail(h, SC_ERR_ARG, "crosslink joins a bond to itself");
  Dev &d = h->d;
  Geo &g = h->g;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  refresh_records(h);
  if ((rc = dalloc(h, &d.xb, 2 * n, false))) return rc;
  if ((rc = dalloc(h, &d.xlam, 2 * n, false))) return rc;
  if ((rc = dalloc(h, &d.xf, 9 * n, false))) return rc;
  if ((rc = dalloc(h, &d.xinc, 2 * n, false))) return rc;
  if ((rc = dalloc(h, &d.xinc_start, nb + 1, false))) return rc;
  if ((rc = dalloc(h, &d.xdeg, nb, false))) return rc;
  if ((rc = dalloc(h, &d.bond_caller, nb, false))) return rc;
  HIPCHK(h, hipMemcpyAsync(d.xb, xb.data(), sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(d.xlam, lambda, sizeof(double) * 2 * n, hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMemcpyAsync(d.bond_caller, caller.data(), sizeof(int32_t) * nb,
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(d.xf, 0, sizeof(double) * 9 * n, h->stream));
  HIPCHK(h, hipMemsetAsync(d.xdeg, 0, sizeof(int32_t) * nb, h->stream));
  g.nx = n;
  g.xk = p->k_spring;
  g.xr0 = p->rest_length;
  g.xkc = p->k_compress;
  g.xpoff = p->k_off > 0 ? 1.0 - std::exp(-p->k_off * h->p.delta) : 0.0;
  g.xrb2 = p->r_bind * p->r_bind;
  g.xseed = p->seed;
  const double lmax = h->max_bond_length;
  const double cell_len = 2.0 * h->p.system_radius / g.n1d;
  g.xreach = (int)std::ceil((p->r_bind + 0.5 * lmax) / cell_len) + 1;
  // build the per-bond lists once
  Ctl c = read_ctl(h);
  c.xl_dirty = 1;
  if ((rc = write_ctl(h, c))) return rc;
  sc::launch_xlink_lists(d, g, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  *h->ctl_host = read_ctl(h);
  drop_graph(h);  // kernels take the system constants by value
  return SC_OK;
}

int sc_get_crosslinks(sc_handle h, int32_t *bonds, double *lambda, double *force,
                      int64_t *events) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t n = h->g.nx;
  if (events) *events = read_ctl(h).xl_events;
  if (n == 0) return SC_OK;
  if (bonds) {
    std::vector<int32_t> xb(2 * n);
    HIPCHK(h, hipMemcpy(xb.data(), h->d.xb, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < 2 * n; ++i)
      bonds[i] = (int32_t)(h->reordered ? h->bond_map[xb[i]] : xb[i]);
  }
  if (lambda)
    HIPCHK(h, hipMemcpy(lambda, h->d.xlam, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
  if (force) HIPCHK(h, hipMemcpy(force, h->d.xf, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
  return SC_OK;
}

int sc_step(sc_handle h, int32_t n_steps) {
  int rc = sc_step_async(h, n_steps);
  if (rc) return rc;
  for (;;) {
    HIPCHK(h, hipEventSynchronize(h->done_ev));
    int done = 0;
    if ((rc = sc_step_poll(h, &done))) return rc;
    if (done) return SC_OK;
  }
}

int sc_step_async(sc_handle h, int32_t n_steps) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (n_steps < 0) return fail(h, SC_ERR_ARG, "negative step count");
  if (!h->pending) {
    if (h->ctl_host->abort)
      return fail(h, h->ctl_host->err ? h->ctl_host->err : SC_ERR_STATE,
                  "handle is in an error state");
    h->target = h->i_step_host;
    h->grow_rounds = 0;
  }
  if (n_steps == 0) return SC_OK;
  const int64_t from = h->target;
  if (h->g.xi_buf && !midsteps_in_buffer(h, from, n_steps))
    return fail(h, SC_ERR_STATE,
                "the caller's noise buffer does not hold these midsteps (sc_set_noise)");
  h->target += n_steps;
  return enqueue_range(h, from, n_steps);
}

int sc_step_poll(sc_handle h, int *done) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (done) *done = 0;
  if (!h->pending) {
    if (done) *done = 1;
    return SC_OK;
  }
  const hipError_t q = hipEventQuery(h->done_ev);
  if (q == hipErrorNotReady) return SC_OK;
  if (q != hipSuccess)
    return fail(h, SC_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
  HIPCHK(h, hipGetLastError());
  Ctl c = *h->ctl_host;
  h->i_step_host = c.i_step;
  refresh_records(h);
  if (!c.abort) {
    h->pending = false;
    if (done) *done = 1;
    return SC_OK;
  }
  if (c.overflow && h->grow_rounds < 32) {
    // the step that overflowed left the state untouched (every kernel after
    // the capacity check returned): grow the candidate buffers and run the
    // remaining steps again from there
    ++h->grow_rounds;
    if ((rc = alloc_cand(h, grown_capacity(h->g.cap, c.nc_need)))) {
      h->pending = false;
      return rc;
    }
    c.abort = 0;
    c.overflow = 0;
    if ((rc = write_ctl(h, c))) {
      h->pending = false;
      return rc;
    }
    return enqueue_range(h, c.i_step, h->target - c.i_step);
  }
  h->pending = false;
  if (c.overflow) return fail(h, SC_ERR_NOMEM, "candidate capacity growth did not converge");
  if (c.err == SC_ERR_UNPHYSICAL)
    return fail(h, SC_ERR_UNPHYSICAL, "Unphysical result: dr > 10000");
  return fail(h, c.err ? c.err : SC_ERR_STATE, "step aborted");
}

int sc_get_sites(sc_handle h, double *r, double *r_prev, double *f_rand) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (r && (rc = fetch_mapped(h, h->d.r, r, h->g.ns, 3, true))) return rc;
  if (r_prev && (rc = fetch_mapped(h, h->d.rp, r_prev, h->g.ns, 3, true))) return rc;
  if (f_rand) {
    if (h->g.compact) {  // mixed precision: FP32 random forces
      const int64_t ns = h->g.ns;
      std::vector<float> tmp(3 * (size_t)ns);
      HIPCHK(h, hipMemcpy(tmp.data(), h->d.fr, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost));
      std::vector<double> wide(tmp.begin(), tmp.end());
      to_caller(h, h->site_map, wide.data(), f_rand, ns, 3);
    } else if ((rc = fetch_mapped(h, h->d.fr, f_rand, h->g.ns, 3, true))) {
      return rc;
    }
  }
  return SC_OK;
}

int sc_get_bonds(sc_handle h, double *pos, double *spos, double *u, double *len) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  refresh_records(h);
  const int64_t nb = h->g.nb, ns = h->g.ns;
  std::vector<double> sp(3 * (size_t)nb), uu(3 * (size_t)nb), ln(nb);
  if (h->g.compact) {
    // mixed precision: the 32-byte record {cell offset x, y, z (2^-32
    // cell), length high}, {octahedral unit vector, packed cell
    // coordinates, length low}, decoded as the kernels decode it
    std::vector<uint32_t> rec(8 * (size_t)nb);
    HIPCHK(h, hipMemcpy(rec.data(), h->d.br, sizeof(uint32_t) * rec.size(), hipMemcpyDeviceToHost));
    for (int64_t b = 0; b < nb; ++b) {
      const uint32_t *w = &rec[8 * b];
      const int packed = (int)w[6];
      const int c[3] = {packed & 1023, (packed >> 10) & 1023, (packed >> 20) & 1023};
      double u[3];
      sc::oct_decode(w[4], w[5], u);
      for (int k = 0; k < 3; ++k) {
        sp[k * nb + b] = ((double)c[k] + (double)w[k] * (1.0 / 4294967296.0)) * h->g.inv_n1d - 0.5;
        uu[k * nb + b] = u[k];
      }
      float lh, ll;
      std::memcpy(&lh, &w[3], 4);
      std::memcpy(&ll, &w[7], 4);
      ln[b] = (double)lh + (double)ll;
    }
  } else {
    std::vector<double> rec(8 * (size_t)nb);
    HIPCHK(h, hipMemcpy(rec.data(), h->d.br, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
    for (int64_t b = 0; b < nb; ++b) {
      for (int k = 0; k < 3; ++k) {
        sp[k * nb + b] = rec[8 * b + k];
        uu[k * nb + b] = rec[8 * b + 3 + k];
      }
      ln[b] = rec[8 * b + 6];
    }
  }
  if (spos) to_caller(h, h->bond_map, sp.data(), spos, nb, 3);
  if (u) to_caller(h, h->bond_map, uu.data(), u, nb, 3);
  if (len) to_caller(h, h->bond_map, ln.data(), len, nb, 1);
  if (pos) {
    // bond centre r1 + 0.5 (r2 - r1) (src/bond.cpp:27-34), from the sites
    std::vector<double> r(3 * (size_t)ns), ps(3 * (size_t)nb);
    HIPCHK(h, hipMemcpy(r.data(), h->d.r, sizeof(double) * r.size(), hipMemcpyDeviceToHost));
    const int nf = h->g.nf;
    for (int f = 0; f < nf; ++f) {
      const int64_t s0 = h->site_off[f], s1 = h->site_off[f + 1];
      for (int64_t i = s0; i < s1 - 1; ++i) {
        const int64_t b = i - f;
        for (int k = 0; k < 3; ++k) {
          const double o = r[k * ns + i + 1] - r[k * ns + i];
          ps[k * nb + b] = r[k * ns + i] + 0.5 * o;
        }
      }
    }
    to_caller(h, h->bond_map, ps.data(), pos, nb, 3);
  }
  return SC_OK;
}

int sc_get_bond_forces(sc_handle h, double *f, double *t) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t nb = h->g.nb;
  const Dev &d = h->d;
  for (int which = 0; which < 2; ++which) {
    double *out = which == 0 ? f : t;
    if (!out) continue;
    if (d.bf) {
      if ((rc = fetch_mapped(h, which == 0 ? d.bf : d.bt, out, nb, 3, false))) return rc;
    } else {  // mixed precision: FP32 sums widened to FP64
      std::vector<float> v32(3 * (size_t)nb);
      HIPCHK(h, hipMemcpy(v32.data(), which == 0 ? d.bf32 : d.bt32, sizeof(float) * v32.size(),
                          hipMemcpyDeviceToHost));
      std::vector<double> v(v32.begin(), v32.end());
      to_caller(h, h->bond_map, v.data(), out, nb, 3);
    }
  }
  return SC_OK;
}

int sc_get_tensions(sc_handle h, double *tension) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!tension) return fail(h, SC_ERR_ARG, "null output");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return fetch_mapped(h, h->d.tension, tension, h->g.nb, 1, false);
}

int sc_thermo_enable(sc_handle h, int on) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int want = on ? 1 : 0;
  if (want && !h->g.thermo)
    HIPCHK(h, hipMemsetAsync(h->d.stress_acc, 0,
                             sizeof(double) * 6 * stress_slots(h->g.nb), h->stream));
  if (want != h->g.thermo) drop_graph(h);  // kernels take the system constants by value
  h->g.thermo = want;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return SC_OK;
}

int sc_calculate_pressure(sc_handle h, int32_t n_thermo, double *pressure_tensor,
                          double *pressure, double *volume) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (n_thermo < 1) return fail(h, SC_ERR_ARG, "need n_thermo >= 1");
  if (!h->g.thermo) return fail(h, SC_ERR_STATE, "thermodynamics not enabled");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int64_t slots = stress_slots(h->g.nb);
  std::vector<double> acc(6 * (size_t)slots);
  HIPCHK(h, hipMemcpy(acc.data(), h->d.stress_acc, sizeof(double) * acc.size(),
                      hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemsetAsync(h->d.stress_acc, 0, sizeof(double) * acc.size(), h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double s6[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t b = 0; b < slots; ++b)
    for (int k = 0; k < 6; ++k) s6[k] += acc[6 * b + k];
  // the pair stress -dr_i F_j is symmetric (F is parallel to dr)
  const double S[9] = {s6[0], s6[3], s6[4], s6[3], s6[1], s6[5], s6[4], s6[5], s6[2]};
  // InteractionManager::CalculatePressure (src/interaction_manager.cpp:
  // 534-559): the bond count over the volume on the diagonal plus the
  // stress over volume and n_thermo; Space::UpdateVolume of the periodic
  // box, the unit cell's determinant (src/space.cpp:233-237)
  const double *hc = h->g.h;
  const double V = hc[0] * (hc[4] * hc[8] - hc[5] * hc[7]) - hc[1] * (hc[3] * hc[8] - hc[5] * hc[6]) +
                   hc[2] * (hc[3] * hc[7] - hc[4] * hc[6]);
  const double inv_V = 1.0 / V;
  double P[9], p = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double v = 0.0;
      if (i == j) v += (double)h->g.nb * inv_V;
      v += S[3 * i + j] * inv_V / n_thermo;
      P[3 * i + j] = v;
    }
  for (int i = 0; i < 3; ++i) p += P[4 * i];
  p /= 3;
  if (pressure_tensor) std::memcpy(pressure_tensor, P, sizeof(P));
  if (pressure) *pressure = p;
  if (volume) *volume = V;
  return SC_OK;
}

int sc_get_bond_energies(sc_handle h, double *pote) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!pote) return fail(h, SC_ERR_ARG, "null output");
  if (!h->g.thermo) return fail(h, SC_ERR_STATE, "thermodynamics not enabled");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return fetch_mapped(h, h->d.bpote, pote, h->g.nb, 1, false);
}

int sc_get_cell_keys(sc_handle h, int32_t *keys) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!keys) return fail(h, SC_ERR_ARG, "null output");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<int32_t> tmp(h->g.nb);
  HIPCHK(h, hipMemcpy(tmp.data(), h->d.keys, sizeof(int32_t) * h->g.nb, hipMemcpyDeviceToHost));
  to_caller(h, h->bond_map, tmp.data(), keys, h->g.nb, 1);
  return SC_OK;
}

int sc_get_candidates(sc_handle h, int32_t *b1, int32_t *b2, size_t cap, size_t *n) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  Ctl c = read_ctl(h);
  if (n) *n = (size_t)c.nc;
  if (!b1 && !b2) return SC_OK;
  if (cap < (size_t)c.nc) return fail(h, SC_ERR_ARG, "output capacity too small");
  std::vector<HostCand> cand;
  if ((rc = host_candidates(h, &cand, nullptr))) return rc;
  for (size_t i = 0; i < cand.size(); ++i) {
    const int64_t x = h->reordered ? h->bond_map[cand[i].b1] : cand[i].b1;
    const int64_t y = h->reordered ? h->bond_map[cand[i].b2] : cand[i].b2;
    if (b1) b1[i] = (int32_t)x;
    if (b2) b2[i] = (int32_t)y;
  }
  return SC_OK;
}

int sc_get_slot_forces(sc_handle h, double *f, size_t cap) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  Ctl c = read_ctl(h);
  if (!f || cap < (size_t)c.nc) return fail(h, SC_ERR_ARG, "output capacity too small");
  const int64_t nc = c.nc, ni = 2 * nc;
  std::vector<HostCand> cand;
  std::vector<int2> inc;
  if ((rc = host_candidates(h, &cand, &inc))) return rc;
  // a candidate's slot: the copy under its obj1 (both copies are equal)
  std::vector<double> v(3 * (size_t)ni, 0.0);
  if (h->p.precision == SC_PREC_FP64) {
    std::vector<double4> s(ni);
    if (ni) HIPCHK(h, hipMemcpy(s.data(), h->d.islot64, sizeof(double4) * ni, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < ni; ++q) {
      v[3 * q] = s[q].x;
      v[3 * q + 1] = s[q].y;
      v[3 * q + 2] = s[q].z;
    }
  } else {
    std::vector<float4> s(ni);
    if (ni) HIPCHK(h, hipMemcpy(s.data(), h->d.islot32, sizeof(float4) * ni, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < ni; ++q) {
      v[3 * q] = s[q].x;
      v[3 * q + 1] = s[q].y;
      v[3 * q + 2] = s[q].z;
    }
  }
  for (int64_t k = 0; k < nc; ++k) {
    const int64_t q = cand[k].q;
    const bool written = (inc[q].y & 2) != 0;  // not written since the rebuild: 0
    f[k] = written ? v[3 * q] : 0.0;
    f[nc + k] = written ? v[3 * q + 1] : 0.0;
    f[2 * nc + k] = written ? v[3 * q + 2] : 0.0;
  }
  return SC_OK;
}

int sc_get_dr_state(sc_handle h, double *acc) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!acc) return fail(h, SC_ERR_ARG, "null output");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return fetch_mapped(h, h->d.acc, acc, h->g.ns, 1, true);
}

int sc_get_counters(sc_handle h, int64_t *i_step, int64_t *n_rebuilds,
                    int64_t *n_candidates, int32_t *n_cells_1d) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  Ctl c = read_ctl(h);
  if (i_step) *i_step = c.i_step;
  if (n_rebuilds) *n_rebuilds = c.n_rebuilds;
  if (n_candidates) *n_candidates = c.nc;
  if (n_cells_1d) *n_cells_1d = h->g.n1d;
  return SC_OK;
}

int sc_set_sites(sc_handle h, const double *r) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!r) return fail(h, SC_ERR_ARG, "null input");
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  refresh_records(h);
  std::vector<double> st;
  if (h->reordered) {
    const int64_t ns = h->g.ns;
    st.resize(3 * (size_t)ns);
    for (int k = 0; k < 3; ++k)
      for (int64_t i = 0; i < ns; ++i) st[k * ns + i] = r[k * ns + h->site_map[i]];
    r = st.data();
  }
  HIPCHK(h, hipMemcpy(h->d.r, r, sizeof(double) * 3 * h->g.ns, hipMemcpyHostToDevice));
  sc::launch_bond_init(h->d, h->g, h->p.precision, h->stream);
  Ctl c = read_ctl(h);
  if (c.i_step >= 2 && c.i_step % 2 == 0) {
    // the last full step already folded its displacement into the rebuild
    // test of the next step; redo that with the new positions
    c.dr_bits = 0ull;
    if ((rc = write_ctl(h, c))) return rc;
    sc::launch_redo_dr(h->d, h->g, h->stream);
  }
  // the decision for the next step was taken from the old positions
  sc::launch_decide(h->d, h->g, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  return SC_OK;
}

int sc_set_noise(sc_handle h, int mode, uint64_t seed, const double *xi, size_t n_xi) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  if (mode != SC_NOISE_COUNTER && mode != SC_NOISE_BUFFER)
    return fail(h, SC_ERR_ARG, "unknown noise mode");
  const size_t per_mid = 3 * (size_t)h->g.ns;
  if (mode == SC_NOISE_BUFFER && (!xi || n_xi == 0 || n_xi % per_mid != 0))
    return fail(h, SC_ERR_ARG, "noise buffer: a positive multiple of 3 * n_sites values");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->xi_dev) {
    (void)hipFree(h->xi_dev);
    h->xi_dev = nullptr;
  }
  h->g.xi_buf = nullptr;
  h->g.xi_m0 = 0;
  h->g.xi_nm = 0;
  if (mode == SC_NOISE_COUNTER) {
    h->g.seed = seed;
  } else {
    double *buf = nullptr;
    if (hipMalloc(&buf, sizeof(double) * n_xi) != hipSuccess)
      return fail(h, SC_ERR_NOMEM, "noise buffer allocation failed");
    h->xi_dev = buf;
    HIPCHK(h, hipMemcpy(buf, xi, sizeof(double) * n_xi, hipMemcpyHostToDevice));
    h->g.xi_buf = buf;
    h->g.xi_m0 = (h->i_step_host + 1) / 2;  // the next midstep's index
    h->g.xi_nm = (int64_t)(n_xi / per_mid);
  }
  drop_graph(h);  // kernels take the system constants by value
  return SC_OK;
}

int sc_filament_stats(sc_handle h, double *e2e, double *com) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!e2e && !com) return SC_OK;
  const int64_t n = 3 * (int64_t)h->g.nf;
  double *buf = nullptr;
  HIPCHK(h, hipMalloc(&buf, sizeof(double) * 2 * n));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  refresh_records(h);
  sc::launch_filament_stats(h->d, h->g, buf, buf + n, h->stream);
  std::vector<double> tmp(2 * (size_t)n);
  hipError_t e = hipMemcpyAsync(tmp.data(), buf, sizeof(double) * 2 * n, hipMemcpyDeviceToHost,
                                h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(h, SC_ERR_HIP, std::string("sc_filament_stats: ") +
                                                      hipGetErrorString(e));
  const int64_t nf = h->g.nf;
  for (int k = 0; k < 3; ++k)
    for (int64_t f = 0; f < nf; ++f) {
      const int64_t u = h->reordered ? h->perm[f] : f;
      if (e2e) e2e[k * nf + u] = tmp[k * nf + f];
      if (com) com[k * nf + u] = tmp[n + k * nf + f];
    }
  return SC_OK;
}

int sc_theta_histogram(sc_handle h, int32_t n_bins, int32_t n_rows, int64_t *hist) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (n_bins <= 0 || n_rows <= 0 || !hist) return fail(h, SC_ERR_ARG, "bad histogram shape");
  const size_t n = (size_t)n_bins * (size_t)n_rows;
  unsigned long long *buf = nullptr;
  HIPCHK(h, hipMalloc(&buf, sizeof(unsigned long long) * n));
  std::vector<unsigned long long> host(n);
  hipError_t e = hipMemsetAsync(buf, 0, sizeof(unsigned long long) * n, h->stream);
  if (e == hipSuccess) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    refresh_records(h);
    sc::launch_theta_hist(h->d, h->g, n_bins, n_rows, buf, h->stream);
    e = hipMemcpyAsync(host.data(), buf, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost,
                       h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(h, SC_ERR_HIP, std::string("sc_theta_histogram: ") +
                                                      hipGetErrorString(e));
  for (size_t k = 0; k < n; ++k) hist[k] += (int64_t)host[k];
  return SC_OK;
}

int sc_theta_histogram_noalloc(sc_handle h, int32_t n_bins, int32_t n_rows, int64_t *hist) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (n_bins <= 0 || n_rows <= 0 || !hist) return fail(h, SC_ERR_ARG, "bad histogram shape");
  const int64_t n = (int64_t)n_bins * (int64_t)n_rows;
  if (n > kScratchCounters)
    return fail(h, SC_ERR_ARG, "histogram of " + std::to_string(n) + " counters exceeds the " +
                                   std::to_string(kScratchCounters) +
                                   "-counter scratch of sc_theta_histogram_noalloc");
  unsigned long long *buf = scratch_ptr(h->device);
  if (!buf) return fail(h, SC_ERR_STATE, "no scratch address for this device");
  // the array keeps the last call's counts: clear the part this call uses
  // (a fill kernel, not a memset node, so the call may be captured)
  HIPCHK(h, hipStreamSynchronize(h->stream));
  refresh_records(h);
  sc::launch_zero_i32(reinterpret_cast<int32_t *>(buf), 2 * n, h->stream);
  sc::launch_theta_hist(h->d, h->g, n_bins, n_rows, buf, h->stream);
  HIPCHK(h, hipGetLastError());
  // counts come back through a fixed host block (no allocation either side)
  constexpr int64_t kChunk = 512;
  unsigned long long host[kChunk];
  for (int64_t k0 = 0; k0 < n; k0 += kChunk) {
    const int64_t m = std::min(kChunk, n - k0);
    HIPCHK(h, hipMemcpyAsync(host, buf + k0, sizeof(unsigned long long) * m,
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int64_t k = 0; k < m; ++k) hist[k0 + k] += (int64_t)host[k];
  }
  return SC_OK;
}

}  // extern "C"

namespace {

struct CkHeader {
  char magic[8];
  int32_t version, precision, spatial_order, nf;
  int64_t ns, nb, nc, n_cells, nx;
  Ctl ctl;
};

constexpr char kCkMagic[8] = {'S', 'C', 'C', 'K', 'P', 'T', '0', '3'};

// One field of the checkpoint: `n` elements of `T` at device address `dev`.
struct CkField {
  void *dev;
  size_t bytes;
};

std::vector<CkField> ck_fields(sc_system *h, int64_t nc) {
  const Dev &d = h->d;
  const int64_t ns = h->g.ns, nb = h->g.nb, nf = h->g.nf;
  const size_t slot = h->p.precision == SC_PREC_FP64 ? sizeof(double4) : sizeof(float4);
  void *slots = h->p.precision == SC_PREC_FP64 ? (void *)d.islot64 : (void *)d.islot32;
  return {{d.r, 24 * (size_t)ns},        {d.rp, 24 * (size_t)ns},
          {d.fr, (h->g.compact ? 12 : 24) * (size_t)ns}, {d.dr0, 24 * (size_t)ns},
          {d.acc, 8 * (size_t)ns},       {d.acc_prev, 8 * (size_t)ns},
          {d.br, (h->g.compact ? 32 : 64) * (size_t)nb},
          {d.bf ? (void *)d.bf : (void *)d.bf32, (d.bf ? 24 : 12) * (size_t)nb},
          {d.bt ? (void *)d.bt : (void *)d.bt32, (d.bt ? 24 : 12) * (size_t)nb},
          {d.tension, 8 * (size_t)nb},
          {d.keys, 4 * (size_t)nb},      {d.norm, 8 * (size_t)nf},
          {slots, 2 * slot * (size_t)nc},
          {d.inc_start, 4 * (size_t)(nb + 1)}, {d.inc, 16 * (size_t)nc},
          {d.xb, 8 * (size_t)h->g.nx}, {d.xlam, 16 * (size_t)h->g.nx},
          {d.xf, 72 * (size_t)h->g.nx}, {d.xinc, 8 * (size_t)h->g.nx},
          {d.xinc_start, h->g.nx ? 4 * (size_t)(nb + 1) : 0}};
}

size_t ck_size(sc_system *h, int64_t nc) {
  size_t b = sizeof(CkHeader) + 4 * (size_t)(h->g.nf + 1) + 4 * (size_t)h->g.nf;
  for (const CkField &f : ck_fields(h, nc)) b += f.bytes;
  return b;
}

}  // namespace

extern "C" {

int sc_checkpoint_save(sc_handle h, void *buf, size_t cap, size_t *size) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const Ctl c = read_ctl(h);
  h->d.br = h->d.br_out = h->rec[c.i_step & 1];
  const size_t need = ck_size(h, c.nc);
  if (size) *size = need;
  if (!buf) return SC_OK;
  if (cap < need) return fail(h, SC_ERR_ARG, "checkpoint buffer too small");
  char *p = static_cast<char *>(buf);
  CkHeader hd{};
  std::memcpy(hd.magic, kCkMagic, 8);
  hd.version = 3;
  hd.precision = h->p.precision;
  hd.spatial_order = h->reordered ? 1 : 0;
  hd.nf = h->g.nf;
  hd.ns = h->g.ns;
  hd.nb = h->g.nb;
  hd.nc = c.nc;
  hd.n_cells = h->g.n_cells;
  hd.nx = h->g.nx;
  hd.ctl = c;
  std::memcpy(p, &hd, sizeof(hd));
  p += sizeof(hd);
  std::memcpy(p, h->site_off.data(), 4 * (size_t)(h->g.nf + 1));
  p += 4 * (size_t)(h->g.nf + 1);
  for (int f = 0; f < h->g.nf; ++f) {
    const int32_t u = h->reordered ? h->perm[f] : f;
    std::memcpy(p, &u, 4);
    p += 4;
  }
  for (const CkField &f : ck_fields(h, c.nc)) {
    if (f.bytes) HIPCHK(h, hipMemcpy(p, f.dev, f.bytes, hipMemcpyDeviceToHost));
    p += f.bytes;
  }
  return SC_OK;
}

int sc_checkpoint_load(sc_handle h, const void *buf, size_t size) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (!buf || size < sizeof(CkHeader)) return fail(h, SC_ERR_ARG, "checkpoint too short");
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  CkHeader hd;
  std::memcpy(&hd, buf, sizeof(hd));
  if (std::memcmp(hd.magic, kCkMagic, 8) != 0 || hd.version != 3)
    return fail(h, SC_ERR_ARG, "not a checkpoint of this library");
  if (hd.precision != h->p.precision || hd.spatial_order != (h->reordered ? 1 : 0) ||
      hd.nf != h->g.nf || hd.ns != h->g.ns || hd.nb != h->g.nb || hd.n_cells != h->g.n_cells ||
      hd.nx != h->g.nx)
    return fail(h, SC_ERR_ARG, "checkpoint of a different system layout");
  if (size != ck_size(h, hd.nc)) return fail(h, SC_ERR_ARG, "checkpoint size mismatch");
  const char *p = static_cast<const char *>(buf) + sizeof(hd);
  if (std::memcmp(p, h->site_off.data(), 4 * (size_t)(h->g.nf + 1)) != 0)
    return fail(h, SC_ERR_ARG, "checkpoint filament sizes differ");
  p += 4 * (size_t)(h->g.nf + 1);
  for (int f = 0; f < h->g.nf; ++f) {
    int32_t u;
    std::memcpy(&u, p, 4);
    p += 4;
    if (u != (h->reordered ? h->perm[f] : f))
      return fail(h, SC_ERR_ARG, "checkpoint storage order differs");
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (hd.nc > h->g.cap) {
    if ((rc = alloc_cand(h, hd.nc + hd.nc / 4 + 1024))) return rc;
  }
  h->d.br = h->d.br_out = h->rec[hd.ctl.i_step & 1];
  for (const CkField &f : ck_fields(h, hd.nc)) {
    if (f.bytes) HIPCHK(h, hipMemcpy(f.dev, p, f.bytes, hipMemcpyHostToDevice));
    p += f.bytes;
  }
  if ((rc = write_ctl(h, hd.ctl))) return rc;
  // cell membership of the saved state's last rebuild (used between
  // rebuilds by the crosslink partner search)
  sc::launch_cells_from_keys(h->d, h->g, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  *h->ctl_host = hd.ctl;
  h->i_step_host = hd.ctl.i_step;
  return SC_OK;
}

int sc_autotune(sc_handle h, int32_t steps, sc_tune_report *report) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  if (steps < 2) steps = 2;
  steps += steps & 1;  // whole midstep/full-step pairs
  // the side-stream shapes (several bins, overlapped ranges) need the
  // side streams
  if ((rc = ensure_side_streams(h))) return rc;
  size_t n = 0;
  if ((rc = sc_checkpoint_save(h, nullptr, 0, &n))) return rc;
  std::vector<char> blob(n);
  if ((rc = sc_checkpoint_save(h, blob.data(), n, &n))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  const bool prof = h->prof;
  h->prof = false;
  // trials step past the caller's noise buffer: they draw from the counter
  // stream instead (their trajectories are discarded)
  const double *xi_keep = h->g.xi_buf;
  h->g.xi_buf = nullptr;
  // nor do they add to the stress sum or the bond energies
  const int thermo_keep = h->g.thermo;
  h->g.thermo = 0;
  // time `steps` steps of configuration t from the saved state
  auto trial = [&](const sc::Tune &t, double *ms) -> int {
    h->tune = t;
    drop_graph(h);
    int r = sc_checkpoint_load(h, blob.data(), blob.size());
    if (r) return r;
    if (h->p.use_graph && (r = capture_graphs(h))) return r;
    // one untimed pair of steps (first launches of fresh graphs)
    if ((r = enqueue_range(h, h->i_step_host, 2))) return r;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->pending = false;
    HIPCHK(h, hipEventRecord(e0, h->stream));
    if ((r = enqueue_range(h, h->i_step_host + 2, steps))) return r;
    HIPCHK(h, hipEventRecord(e1, h->stream));
    HIPCHK(h, hipEventSynchronize(e1));
    h->pending = false;
    if (h->ctl_host->abort) return fail(h, SC_ERR_STATE, "step aborted during tuning");
    float f = 0.f;
    HIPCHK(h, hipEventElapsedTime(&f, e0, e1));
    *ms = f / steps;
    return SC_OK;
  };
  const sc::Tune start = h->tune;
  double best_ms = 0, start_ms = 0;
  sc::Tune best = start;
  int trials = 0;
  std::vector<std::pair<double, sc::Tune>> seen;  // every shape timed, and its time
  rc = trial(start, &start_ms);
  best_ms = start_ms;
  if (rc == SC_OK) seen.push_back({start_ms, start});
  // Two stages: the pair kernel's shape with the integrator's fixed, then
  // the integrator's with the best pair shape (the two kernels are
  // separate launches, so their times add).
  auto consider = [&](const sc::Tune &t) {
    if (rc) return;
    double ms = 0;
    rc = trial(t, &ms);
    ++trials;
    if (rc == SC_OK) seen.push_back({ms, t});
    if (rc == SC_OK && ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  };
  // pair kernel: per-bond with 128/256/512 threads per block, or the
  // balanced one-wave-per-64-bonds variant (block size not used)
  const int blocks[] = {128, 256, 512, 0};
  for (int bs : blocks)
    for (int x = 0; x < 2; ++x) {
      sc::Tune t = best;
      t.pair_bal = bs == 0 ? 1 : 0;
      t.pair_block = bs == 0 ? 256 : bs;
      t.pair_xcd = x;
      consider(t);
    }
  // integrator: waves per block and, with several launches, launches on
  // side streams; on the packed layout also one launch for all lane-group
  // widths or one per width (the same arithmetic).  Sites per lane (layouts
  // 0 and 1) change the rounding of the mixed path and are not tuned.
  const sc::Tune pair_best = best;
  const int waves[] = {1, 2, 4};
  const bool packed = h->tune.int_layout >= 2 && h->pack_waves > 0;
  std::vector<int> layouts = {h->tune.int_layout};
  if (packed) layouts = {2, 3};
  for (int lay : layouts) {
    // side streams only matter with several integrator launches (0..3 of
    // them); layout 2 runs the 256-site bin inside its one launch
    int launches = 0;
    for (const Bin &b : h->bins)
      if (!packed || (b.maxn > 128 && !(lay == 2 && b.maxn == 256))) ++launches;
    if (packed && lay == 2) ++launches;
    if (packed && lay == 3)
      for (int c = 0; c < 3; ++c) launches += h->pack_cs[c + 1] > h->pack_cs[c] ? 1 : 0;
    const int n_str = std::min(launches, 1 + sc_system::kSide);
    for (int w : waves)
      for (int sm = 0; sm < n_str; ++sm) {
        sc::Tune t = pair_best;
        t.int_waves = w;
        t.int_streams = sm;
        t.int_layout = lay;
        consider(t);
      }
  }
  // overlap of the pair pass and the integrator over bond ranges (systems
  // without crosslinks); the ranges take the side stream, so the
  // integrator's own side-stream spreading is off there
  if (h->g.nx == 0 && h->side[0]) {
    const sc::Tune int_best = best;
    const int parts[] = {2, 4, 8};
    for (int pp : parts) {
      sc::Tune t = int_best;
      t.parts = pp;
      t.int_streams = 0;
      consider(t);
    }
  }
  // early integration of the long filaments beside the pair pass
  if (h->g.nx == 0 && h->side[1] && h->long_last >= 0) {
    sc::Tune t = best;
    t.early_long = 1;
    t.parts = 1;
    consider(t);
  }
  // the four fastest shapes are timed twice more, interleaved, and each
  // scored by the fastest of its three timings (a trial is only ever
  // slowed by noise)
  if (rc == SC_OK && seen.size() > 1) {
    auto same = [](const sc::Tune &a, const sc::Tune &b) {
      return a.pair_block == b.pair_block && a.pair_xcd == b.pair_xcd &&
             a.pair_bal == b.pair_bal && a.int_waves == b.int_waves &&
             a.int_layout == b.int_layout && a.int_streams == b.int_streams &&
             a.parts == b.parts && a.early_long == b.early_long;
    };
    std::stable_sort(seen.begin(), seen.end(),
                     [](const std::pair<double, sc::Tune> &a,
                        const std::pair<double, sc::Tune> &b) { return a.first < b.first; });
    std::vector<std::pair<double, sc::Tune>> top;
    for (const auto &e : seen) {
      bool dup = false;
      for (const auto &u : top) dup = dup || same(u.second, e.second);
      if (!dup) top.push_back(e);
      if (top.size() == 4) break;
    }
    for (int round = 0; round < 2; ++round)
      for (auto &e : top) {
        if (rc) break;
        double ms = 0;
        rc = trial(e.second, &ms);
        ++trials;
        if (rc == SC_OK && ms < e.first) e.first = ms;
      }
    if (rc == SC_OK) {
      const auto fastest = std::min_element(
          top.begin(), top.end(),
          [](const std::pair<double, sc::Tune> &a, const std::pair<double, sc::Tune> &b) {
            return a.first < b.first;
          });
      best = fastest->second;
      best_ms = fastest->first;
    }
  }
  // leave the system where it was, with the chosen shapes
  h->tune = rc ? start : best;
  drop_graph(h);
  h->g.xi_buf = xi_keep;
  h->g.thermo = thermo_keep;
  const int rc2 = sc_checkpoint_load(h, blob.data(), blob.size());
  h->prof = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (rc2) return rc2;
  if (report) {
    (void)sc_get_launch_shape(h, &report->shape);
    report->n_trials = trials;
    report->step_ms_start = start_ms;
    report->step_ms_best = best_ms;
  }
  return SC_OK;
}

int sc_set_launch_shape(sc_handle h, const sc_launch_shape *shape) {
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (!shape) return fail(h, SC_ERR_ARG, "null launch shape");
  const sc_launch_shape &v = *shape;
  if ((v.pair_block != 0 && v.pair_block != 64 && v.pair_block != 128 && v.pair_block != 256 &&
       v.pair_block != 512) ||
      (v.pair_xcd != 0 && v.pair_xcd != 1) ||
      (v.integrate_waves != 1 && v.integrate_waves != 2 && v.integrate_waves != 4) ||
      (v.integrate_layout < 0 || v.integrate_layout > 3) ||
      (v.integrate_streams < 0 || v.integrate_streams > 3) ||
      (v.pair_parts < 0 || v.pair_parts > sc::kMaxParts) ||
      (v.early_long != 0 && v.early_long != 1))
    return fail(h, SC_ERR_ARG, "unsupported launch shape");
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  h->tune.pair_bal = v.pair_block == 0 ? 1 : 0;
  h->tune.pair_block = v.pair_block == 0 ? 256 : v.pair_block;
  h->tune.pair_xcd = v.pair_xcd;
  h->tune.int_waves = v.integrate_waves;
  h->tune.int_layout = v.integrate_layout;
  h->tune.int_streams = v.integrate_streams;
  h->tune.parts = v.pair_parts > 1 ? v.pair_parts : 1;
  h->tune.early_long = v.early_long;
  drop_graph(h);
  if (v.integrate_streams > 0 || v.pair_parts > 1 || v.early_long) return ensure_side_streams(h);
  return SC_OK;
}

int sc_get_launch_shape(sc_handle h, sc_launch_shape *shape) {
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (!shape) return fail(h, SC_ERR_ARG, "null launch shape");
  const sc::Tune &t = h->tune;
  shape->pair_block = t.pair_bal ? 0 : t.pair_block;
  shape->pair_xcd = t.pair_xcd;
  shape->integrate_waves = t.int_waves;
  shape->integrate_layout = t.int_layout;
  shape->integrate_streams = t.int_streams;
  shape->pair_parts = t.parts;
  shape->early_long = t.early_long;
  return SC_OK;
}

int sc_device_bytes(sc_handle h, size_t *bytes) {
  if (!h || !bytes) return SC_ERR_ARG;
  *bytes = h->fixed_bytes + h->cand_bytes;
  return SC_OK;
}

int sc_profile_enable(sc_handle h, int on) {
  if (!h) return SC_ERR_ARG;
  if (on && !h->ev[0])
    for (auto &e : h->ev) HIPCHK(h, hipEventCreate(&e));
  h->prof = on != 0;
  for (int k = 0; k < kFamilies; ++k) {
    h->prof_ms[k] = 0;
    h->prof_n[k] = 0;
  }
  return SC_OK;
}

int sc_profile_read(sc_handle h, double *avg_ms, int64_t *launches, int n) {
  if (!h) return SC_ERR_ARG;
  for (int k = 0; k < n && k < kFamilies; ++k) {
    if (avg_ms) avg_ms[k] = h->prof_n[k] ? h->prof_ms[k] / h->prof_n[k] : 0.0;
    if (launches) launches[k] = h->prof_n[k];
  }
  return SC_OK;
}

int sc_time_pair(sc_handle h, int32_t launches, double *avg_ms) {
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (launches < 1 || !avg_ms) return fail(h, SC_ERR_ARG, "need launches >= 1 and an output");
  if (h->pending) return fail(h, SC_ERR_STATE, "steps in flight: poll them first");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(h, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(h, SC_ERR_HIP, "hipEventCreate");
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return fail(h, SC_ERR_HIP, "hipStreamSynchronize");
  }
  refresh_records(h);
  // the pass rewrites the same slot values and sums on an unchanged state,
  // so repeating it leaves the system as it was
  sc::launch_pair(h->d, h->g, h->p.precision, h->tune, h->stream, false);  // warm
  (void)hipEventRecord(e0, h->stream);
  for (int k = 0; k < launches; ++k)
    sc::launch_pair(h->d, h->g, h->p.precision, h->tune, h->stream, false);
  (void)hipEventRecord(e1, h->stream);
  const hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(h, SC_ERR_HIP, hipGetErrorString(e));
  *avg_ms = ms / launches;
  return SC_OK;
}

int sc_copy_probe(size_t bytes, int repeats, double *avg_ms) {
  if (bytes < 32 || repeats <= 0) return SC_ERR_ARG;
  const int64_t n32 = (int64_t)(bytes / 32);
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n32 * 32) != hipSuccess) return SC_ERR_NOMEM;
  if (hipMalloc(&b, n32 * 32) != hipSuccess) {
    (void)hipFree(a);
    return SC_ERR_NOMEM;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SC_OK;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMemsetAsync(a, 1, n32 * 32, st) != hipSuccess) {
    rc = SC_ERR_HIP;
  } else {
    sc::launch_copy_probe(a, b, n32, st);  // warm
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < repeats; ++r) sc::launch_copy_probe(a, b, n32, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) rc = SC_ERR_HIP;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (avg_ms) *avg_ms = ms / repeats;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

const char *sc_last_error(sc_handle h) {
  if (!h) return "null handle";
  return h->err.c_str();
}

int sc_destroy(sc_handle h) {
  if (!h) return SC_ERR_ARG;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  free_cand(h);
  for (void *v : h->fixed_allocs) dfree(h, v);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->done_ev) (void)hipEventDestroy(h->done_ev);
  if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
  for (int k = 0; k < sc_system::kSide; ++k) {
    if (h->side[k]) {
      (void)hipStreamSynchronize(h->side[k]);
      (void)hipStreamDestroy(h->side[k]);
    }
    if (h->join_ev[k]) (void)hipEventDestroy(h->join_ev[k]);
  }
  for (int k = 0; k < sc::kMaxParts; ++k)
    if (h->part_ev[k]) (void)hipEventDestroy(h->part_ev[k]);
  if (h->ctl_host && h->own_ctl_host) (void)hipHostFree(h->ctl_host);
  if (h->xi_dev) (void)hipFree(h->xi_dev);
  if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SC_OK;
}

}  // extern "C"
