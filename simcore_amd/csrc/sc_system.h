// sc_system.h -- pieces of the system handle shared between the single-system
// ABI (sc_api.cpp) and the ensemble runner (sc_ensemble.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <unordered_map>

#include "../../include/simcore_amd.h"

namespace sc {

// First-fit sub-allocator over one device allocation.  Systems admitted by
// the ensemble runner take all their buffers from it, so admission and
// retirement never call hipMalloc/hipFree (hipFree waits for the whole
// device) and the runner's footprint is fixed at creation.
class Arena {
 public:
  static constexpr size_t kAlign = 256;
  static size_t round(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

  int init(size_t bytes);  // SC_OK / SC_ERR_NOMEM
  void release();
  ~Arena() { release(); }

  void *get(size_t bytes);  // nullptr when no free block fits
  void put(void *p);
  size_t capacity() const { return cap_; }
  size_t used() const { return used_bytes_; }
  size_t largest_free() const;

 private:
  char *base_ = nullptr;
  size_t cap_ = 0;
  size_t used_bytes_ = 0;
  std::map<size_t, size_t> free_;             // offset -> size, address order
  std::unordered_map<size_t, size_t> taken_;  // offset -> size
};

// sc_create with a caller-owned stream (not destroyed with the handle), an
// optional arena for every device buffer and an optional caller-owned
// pinned control-block mirror (pinned allocations and frees are slow and
// may wait for the device; the ensemble runner keeps one per slot).
struct Ctl;
int system_create(const sc_params *params, hipStream_t stream, Arena *arena,
                  sc_handle *out, Ctl *pinned_ctl = nullptr);

// Ensemble batches (grouped launches, sc_internal.h).  A system can join a
// batch when it is loaded, idle, at an even step count (batches take an odd
// and an even step in turn), in mixed precision with every filament in the
// packed waves or the 256-site bin, and without crosslinks, thermodynamic
// sums, a caller noise buffer or profiling.
struct Member;
bool system_batchable(sc_handle h);
// The system's entry in a batch's member table.
void system_member(sc_handle h, Member *m);
// After a batch took the system through step `target`: its control block
// as the batch left it.  Same outcomes as sc_step_poll: SC_OK with *done
// set, a capacity overflow re-run on the system's own stream (*done 0:
// poll it), or the step's error.
int system_batch_done(sc_handle h, const Ctl &c, int64_t target, int *done);

// Device bytes sc_load_filaments allocates for a system of nf filaments and
// ns sites under these parameters (before any candidate-capacity growth).
size_t system_estimate_bytes(const sc_params *params, int32_t nf, int64_t ns,
                             int n_bins);

}  // namespace sc
