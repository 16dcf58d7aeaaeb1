// sc_kernels.hip -- gfx950 kernels of the filament BD step.
//
// Built with -ffp-contract=off: every FP64 multiply and add rounds on its
// own, in the order the reference's source states it, so the FP64 path
// reproduces the CPU path value-for-value.  Division and sqrt on FP64 are
// IEEE correctly rounded on gfx950 (v_div_scale/fmas/fixup sequence and
// the refined v_sqrt_f64 expansion).
//
// One step, in stream order (all guarded by the device control block,
// which holds the decision for the step about to run):
//   [rebuild: keys (+ column populations), column scan, column scatter,
//    column rank sort (+ cell runs, occupancy), degrees, degree scan,
//    incidence lists (+ capacity, zero dr)]
//   -> pair+scatter -> integrate (per length class)
//   -> commit (and decide the next step)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

// ---------------------------------------------------------------------------
// Control kernels
// ---------------------------------------------------------------------------

// Replaces InteractionManager::CheckUpdateInteractions / GetDrMax
// (src/interaction_manager.cpp:284-312) for the step after i_step: the
// accumulator of a full step is compared with 0.25*cell_len^2 at the next
// (odd) step; the mesh skips the update after a midstep
// (src/mesh.cpp:533-536).  A full step clears the accumulator it is about
// to fill.
__device__ __forceinline__ void decide_next(Ctl *c, double thr) {
  const int64_t k = c->i_step + 1;
  c->cur_step = k;
  c->midstep = (k & 1) ? 1 : 0;
  int rb = 0;
  if ((k & 1) && k >= 3) {
    const double drm = __longlong_as_double((long long)c->dr_bits);
    if (drm > 10000) {  // include/simcore/species.hpp:114-118
      c->abort = 1;
      c->err = SC_ERR_UNPHYSICAL;
      return;
    }
    if (drm > thr) rb = 1;
  }
  if (!(k & 1)) c->dr_bits = 0ull;
  c->rebuild = rb;
}

// Folds the per-wave displacement maxima into Ctl::dr_bits and clears the
// slots; one 64-lane block (the launch of k_decide / k_commit).
__device__ __forceinline__ void fold_dr_slots(Ctl *c, unsigned long long *slots) {
  const int lane = threadIdx.x;
  unsigned long long m = 0ull;
  for (int i = lane; i < kDrSlots; i += 64) {
    const unsigned long long v = slots[i];
    m = v > m ? v : m;
    slots[i] = 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(m, o, 64);
    m = y > m ? y : m;
  }
  if (lane == 0 && m > c->dr_bits) c->dr_bits = m;
}

// standalone decision: after loading or after the host changed the state
__global__ void k_decide(Ctl *c, unsigned long long *slots, double thr) {
  if (blockIdx.x != 0 || c->abort) return;
  fold_dr_slots(c, slots);
  if (threadIdx.x == 0) decide_next(c, thr);
}

// end of a step: commit it and decide the next one (one launch for both).
// Every value it needs is loaded up front, so the kernel costs one memory
// round trip; the decision is decide_next's on those values.
__device__ __forceinline__ void commit_body(Ctl *c, unsigned long long *slots, double thr) {
  static_assert(kDrSlots == 4 * 64, "four slots per lane");
  const int lane = threadIdx.x;
  unsigned long long v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = slots[lane + 64 * i];
  const int32_t abort = c->abort;
  const int64_t cur = c->cur_step;
  const unsigned long long drb = c->dr_bits;
  if (abort) return;
  unsigned long long m = drb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    m = v[i] > m ? v[i] : m;
    slots[lane + 64 * i] = 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(m, o, 64);
    m = y > m ? y : m;
  }
  if (lane != 0) return;
  const int64_t k = cur + 1;  // the next step
  c->i_step = cur;
  c->cur_step = k;
  c->midstep = (k & 1) ? 1 : 0;
  int rb = 0;
  if ((k & 1) && k >= 3) {
    const double drm = __longlong_as_double((long long)m);
    if (drm > 10000) {  // include/simcore/species.hpp:114-118
      c->dr_bits = m;
      c->abort = 1;
      c->err = SC_ERR_UNPHYSICAL;
      return;
    }
    if (drm > thr) rb = 1;
  }
  c->dr_bits = (k & 1) ? m : 0ull;  // a full step clears what it will fill
  c->rebuild = rb;
}

__global__ void k_commit(Ctl *c, unsigned long long *slots, double thr) {
  if (blockIdx.x != 0) return;
  commit_body(c, slots, thr);
}

// one block per member
__global__ void k_group_commit(const Member *__restrict__ ms) {
  const Member &M = ms[blockIdx.x];
  commit_body(M.d[0].ctl, M.d[0].dr_slots, M.g.thr);
}

__global__ void k_group_ctl(const Member *__restrict__ ms, int n, Ctl *out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) out[m] = *ms[m].d[0].ctl;
}

__global__ void k_zero_i32(int32_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = 0;
}

__device__ __forceinline__ bool guard_off(const Ctl *c, int guard) {
  if (!c) return false;
  if (c->abort) return true;
  if (guard == kGuardXlink) return !c->xl_dirty;
  return guard == kGuardRebuild && !c->rebuild;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (int32), three passes.  out has n+1 entries,
// out[n] = total.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int block_excl_scan256(int v, int *total,
                                                  int *wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanBlock / 64; ++i) {
    int s = wsum[i];
    if (i < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__device__ __forceinline__ void scan_reduce_body(const int32_t *in, int64_t n, int32_t *bsum,
                                                 int64_t bid, int64_t nblk) {
  __shared__ int wsum[kScanBlock / 64];
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  for (int64_t t = bid; t < ntiles; t += nblk) {
    const int64_t base = t * kScanTile;
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      const int64_t i = base + (int64_t)j * kScanBlock + threadIdx.x;
      if (i < n) s += in[i];
    }
    int tot;
    block_excl_scan256(s, &tot, wsum);
    if (threadIdx.x == 0) bsum[t] = tot;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(
    const int32_t *in, int64_t n, int32_t *bsum, const Ctl *ctl, int guard) {
  if (guard_off(ctl, guard)) return;
  scan_reduce_body(in, n, bsum, blockIdx.x, gridDim.x);
}

__device__ void tile_scan(int *tile, int *wsum, int carry_in, int *total) {
  // tile holds kScanTile ints; exclusive scan in place plus carry_in.
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = tile[threadIdx.x * kScanItems + j];
    s += v[j];
  }
  int tot;
  int ex = block_excl_scan256(s, &tot, wsum) + carry_in;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    tile[threadIdx.x * kScanItems + j] = ex;
    ex += v[j];
  }
  *total = tot;
  __syncthreads();
}

// Second pass: each tile adds up the sums of the tiles before it (at most a
// few thousand, read from L2) and scans itself; the last tile also writes
// the total.  Two launches per scan; blocks loop over tiles, so any grid
// size serves.
__device__ __forceinline__ void scan_down_body(const int32_t *in, int64_t n, const int32_t *bsum,
                                               int32_t *out, int64_t bid, int64_t nblk) {
  __shared__ int tile[kScanTile];
  __shared__ int wsum[kScanBlock / 64];
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  for (int64_t t = bid; t < ntiles; t += nblk) {
    int part = 0;
    for (int64_t i = threadIdx.x; i < t; i += kScanBlock) part += bsum[i];
    int carry;
    block_excl_scan256(part, &carry, wsum);
    const int64_t base = t * kScanTile;
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
      tile[j] = (base + j < n) ? in[base + j] : 0;
    __syncthreads();
    int tot;
    tile_scan(tile, wsum, carry, &tot);
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
      if (base + j < n) out[base + j] = tile[j];
    if (t == ntiles - 1 && threadIdx.x == 0) out[n] = carry + tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const int32_t *in,
                                                          int64_t n,
                                                          const int32_t *bsum,
                                                          int32_t *out,
                                                          const Ctl *ctl,
                                                          int guard) {
  if (guard_off(ctl, guard)) return;
  scan_down_body(in, n, bsum, out, blockIdx.x, gridDim.x);
}

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *tmp,
                 const Ctl *ctl, int guard, hipStream_t st) {
  int64_t nblk = (n + kScanTile - 1) / kScanTile;
  if (nblk == 0) nblk = 1;
  if (nblk > kStreamBlocks) nblk = kStreamBlocks;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblk), dim3(kScanBlock), 0,
                     st, in, n, tmp, ctl, guard);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nblk), dim3(kScanBlock), 0,
                     st, in, n, tmp, out, ctl, guard);
}

// ---------------------------------------------------------------------------
// Rebuild stage
// ---------------------------------------------------------------------------

// CellList::FindCellCoords over every bond (src/cell_list.cpp:139-157) and
// the population of each cell column (x, y) (n1d^2 counters instead of
// n1d^3, so the scan that follows is small), with each bond's place in its
// column.  Streaming passes loop over their range, so any grid size serves.
// The column counts are aggregated per block tile first: each bond enters
// its column into an LDS hash table (linear probing; a tile of kScanBlock
// bonds has at most that many columns, so it always fits) and takes its
// place among the tile's bonds of that column from an LDS counter; one
// device-scope atomic per distinct column of the tile then hands out the
// tile's block of places.  With spatially ordered storage a tile touches
// a few tens of columns, so the device-scope atomics (which on MI355X go
// past the XCD's L2 to memory) drop by an order of magnitude.  A bond's
// place is any unique slot of its column; the column sort orders them.
__device__ __forceinline__ void keys_body(const Dev &d, const Geo &g, int64_t bid, int64_t nblk) {
  __shared__ int32_t hcol[kScanBlock], hcnt[kScanBlock], hbase[kScanBlock];
  const int t = threadIdx.x;
  const int64_t ntiles = (g.nb + kScanBlock - 1) / kScanBlock;
  for (int64_t tile = bid; tile < ntiles; tile += nblk) {
    hcol[t] = -1;
    hcnt[t] = 0;
    __syncthreads();
    const int64_t b = tile * kScanBlock + t;
    int32_t k = 0, slot = 0, local = 0;
    if (b < g.nb) {
      if (g.compact) {
        k = rec_key_compact(d.br, g, b);
      } else {
        const double *rec = d.br + 8 * b;
        k = cell_key(g.n1d, rec[0], rec[1], rec[2]);
      }
      const int32_t c = k / g.n1d;
      slot = (int32_t)(((uint32_t)c * 2654435761u) >> 24);  // 8 bits of a multiplicative hash
      for (;;) {
        const int32_t prev = atomicCAS(&hcol[slot], -1, c);
        if (prev == -1 || prev == c) break;
        slot = (slot + 1) & (kScanBlock - 1);
      }
      local = atomicAdd(&hcnt[slot], 1);
    }
    __syncthreads();
    if (hcol[t] >= 0) hbase[t] = atomicAdd(&d.col_count[hcol[t]], hcnt[t]);
    __syncthreads();
    if (b < g.nb) {
      d.keys[b] = k;
      d.col_rank[b] = hbase[slot] + local;
    }
    __syncthreads();
  }
}

__global__ void k_keys(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  keys_body(d, g, blockIdx.x, gridDim.x);
}

// Column populations from the stored keys (cell list restore).
__global__ void k_key_count(Dev d, Geo g) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < g.nb; b += stride)
    d.col_rank[b] = atomicAdd(&d.col_count[d.keys[b] / g.n1d], 1);
}

// Groups the bonds by column at the place the population count handed
// each bond (order inside a column arbitrary; the rank sort orders it): a
// pure scatter, no atomics.
__device__ __forceinline__ void scatter_body(const Dev &d, const Geo &g, int64_t bid,
                                             int64_t nblk) {
  const int64_t stride = nblk * blockDim.x;
  for (int64_t b = bid * blockDim.x + threadIdx.x; b < g.nb; b += stride) {
    const int32_t c = d.keys[b] / g.n1d;
    d.col_tmp[d.col_start[c] + d.col_rank[b]] = (int32_t)b;
  }
}

__global__ void k_scatter(Dev d, Geo g, int guard) {
  if (guard_off(d.ctl, guard)) return;
  scatter_body(d, g, blockIdx.x, gridDim.x);
}

// Each cell column sorted by (cell key, bond index): one wavefront per
// column; an entry's place is the number of smaller (key, bond) pairs in
// its column, counted against the column staged in LDS (a column holds a
// few tens of bonds at the densities simulated; longer ones than the stage
// count against global memory).  Members of a cell thus appear in bond
// order, the order Cell::AddObj appends in (src/cell.cpp:5-8).  The pass
// also writes the column's occupancy words (every word of the column, so
// no clearing pass is needed) with each word's count of occupied cells
// before it, and the column's run table: the first sorted entry of each
// occupied cell in z order, at the column's own range of `cell_first`, and
// the number of occupied cells.  The degree pass finds a cell's run from
// its rank among the column's occupied cells (cell_run_of) in these small,
// L2-resident arrays, which cut its HBM reads from 540 to 200 MB on the
// uneven 1.6M-bond system; the cell-indexed runs (n1d^3 entries, mostly
// empty, so every lookup is a random HBM line) serve the list pass's rare
// lookups beyond a bond's run slab (measured faster there) and the
// crosslink partner search.
constexpr int kColStage = 256;

__device__ __forceinline__ int64_t kb_pair(const Dev &d, int32_t b) {
  return ((int64_t)d.keys[b] << 32) | (uint32_t)b;
}

// WPB wavefronts per block, each on its own column (its own LDS stage); the
// blocks step over groups of WPB columns, so every wave of a block runs the
// same number of rounds and meets the same barriers (a wave past the last
// column idles through its round).
template <int WPB>
__device__ __forceinline__ void colsort_body(const Dev &d, const Geo &g, int64_t bid,
                                             int64_t nblk) {
  __shared__ int64_t ks_w[WPB][kColStage], kr_w[WPB][kColStage];
  __shared__ uint32_t occ_w[WPB][64];  // occ_words <= 41 (n1d^3 < 2^31)
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int64_t *ks = ks_w[wv], *kr = kr_w[wv];
  uint32_t *occ = occ_w[wv];
  const int n = g.n1d, W = g.occ_words;
  const int64_t ncol = (int64_t)n * n;
  const int64_t ngrp = (ncol + WPB - 1) / WPB;
  for (int64_t grp = bid; grp < ngrp; grp += nblk) {
    const int64_t c = grp * WPB + wv;
    const bool live = c < ncol;
    const int32_t s = live ? d.col_start[c] : 0, e = live ? d.col_start[c + 1] : 0, m = e - s;
    const bool staged = m <= kColStage;
    if (lane < W) occ[lane] = 0u;
    if (live && lane == 0) d.col_count[c] = 0;  // for the next build's count
    if (staged)
      for (int i = lane; i < m; i += 64) ks[i] = kb_pair(d, d.col_tmp[s + i]);
    __syncthreads();
    for (int i = lane; i < m; i += 64) {
      const int64_t ki = staged ? ks[i] : kb_pair(d, d.col_tmp[s + i]);
      int32_t r = 0;
      if (staged) {
        for (int j = 0; j < m; ++j) r += ks[j] < ki;
        kr[r] = ki;
      } else {
        for (int j = 0; j < m; ++j) r += kb_pair(d, d.col_tmp[s + j]) < ki;
      }
      d.sorted[s + r] = (int32_t)(uint32_t)ki;
      d.skey[s + r] = (int32_t)(ki >> 32);
    }
    // runs: staged columns read the sorted keys from LDS, long ones from
    // global memory after a device-scope fence; each run start's rank among
    // the column's run starts comes from a ballot per 64 entries
    if (!staged) __threadfence();
    __syncthreads();
    int32_t ncell = 0;
    for (int base = 0; base < m; base += 64) {
      const int i = base + lane;
      int32_t k = -1;
      bool start = false;
      if (i < m) {
        k = staged ? (int32_t)(kr[i] >> 32) : d.skey[s + i];
        const int32_t kp =
            i == 0 ? -1 : (staged ? (int32_t)(kr[i - 1] >> 32) : d.skey[s + i - 1]);
        start = k != kp;
      }
      const unsigned long long bal = __ballot(start);
      if (start) {
        const int32_t rank = ncell + __popcll(bal & ((1ull << lane) - 1ull));
        d.cell_first[s + rank] = s + i;
        int32_t j = i + 1;
        while (j < m && (staged ? (int32_t)(kr[j] >> 32) : d.skey[s + j]) == k) ++j;
        d.cell_run[k] = make_int2(s + i, s + j);
        const int z = k - (int32_t)c * n;
        atomicOr(&occ[z >> 5], 1u << (z & 31));
      }
      ncell += __popcll(bal);
    }
    __syncthreads();
    if (live && lane < W) {
      int pre = 0;
      for (int w = 0; w < lane; ++w) pre += __popc(occ[w]);
      d.cell_occ[c * W + lane] = occ[lane];
      d.occ_pre[c * W + lane] = (uint16_t)pre;
    }
    if (live && lane == 0) d.col_ncell[c] = ncell;
    __syncthreads();
  }
}

// Run [start, end) in `sorted` of the occupied cell (x, y, z): its rank
// among its column's occupied cells (the occupancy bits below z plus the
// word's prefix count) indexes the column's run table.
__device__ __forceinline__ int2 cell_run_of(const Dev &d, const Geo &g, int x, int y, int z) {
  const int64_t c = (int64_t)x * g.n1d + y;
  const int64_t wi = c * g.occ_words + (z >> 5);
  const uint32_t word = d.cell_occ[wi];
  const int32_t s0 = d.col_start[c];
  const int32_t nc = d.col_ncell[c];
  const int32_t rank = d.occ_pre[wi] + __popc(word & ((1u << (z & 31)) - 1u));
  const int32_t a = d.cell_first[s0 + rank];
  const int32_t b = rank + 1 < nc ? d.cell_first[s0 + rank + 1] : d.col_start[c + 1];
  return make_int2(a, b);
}

__device__ __forceinline__ int2 cell_run_key(const Dev &d, const Geo &g, int32_t k) {
  const int n = g.n1d;
  return cell_run_of(d, g, k / (n * n), (k / n) % n, k % n);
}


// four columns per 256-thread block: as many waves in flight as one
// column per 64-thread block, a quarter of the workgroups (the launch on
// the odd steps that do not rebuild costs less)
constexpr int kColWaves = 4;

__global__ __launch_bounds__(64 * kColWaves) void k_colsort(Dev d, Geo g, int guard) {
  if (guard_off(d.ctl, guard)) return;
  colsort_body<kColWaves>(d, g, blockIdx.x, gridDim.x);
}

static inline unsigned colsort_grid(int64_t ncol) {
  return (unsigned)std::min<int64_t>((ncol + kColWaves - 1) / kColWaves, 8192 / kColWaves);
}

// Coordinates of the q-th cell of the half-shell stencil of (x, y, z) in
// the order of CellList::AssignCellNeighbors (src/cell_list.cpp:206-262):
// (x+1, y, z); (x-1..x+1, y+1, z); then (x-1..x+1, y-1..y+1, z+1), x
// fastest.  BACK: the mirrored half shell, the 13 cells whose half shell
// holds (x, y, z) (offsets negated).  Periodic wrap; with at least 3 cells
// per side the 13 + 13 + 1 cells are distinct.
template <bool BACK>
__device__ __forceinline__ void shell_cell(int n, int x, int y, int z, int q, int *xo, int *yo,
                                           int *zo) {
  int dx, dy, dz;
  if (q == 0) {
    dx = 1;
    dy = 0;
    dz = 0;
  } else if (q < 4) {
    dx = q - 2;
    dy = 1;
    dz = 0;
  } else {
    dx = (q - 4) % 3 - 1;
    dy = (q - 4) / 3 - 1;
    dz = 1;
  }
  if (BACK) {
    dx = -dx;
    dy = -dy;
    dz = -dz;
  }
  int xp = x + dx, yp = y + dy, zp = z + dz;
  xp = xp < 0 ? n - 1 : (xp == n ? 0 : xp);
  yp = yp < 0 ? n - 1 : (yp == n ? 0 : yp);
  zp = zp < 0 ? n - 1 : (zp == n ? 0 : zp);
  *xo = xp;
  *yo = yp;
  *zo = zp;
}

template <bool BACK>
__device__ __forceinline__ int32_t shell_key(int n, int x, int y, int z, int q) {
  int xp, yp, zp;
  shell_cell<BACK>(n, x, y, z, q, &xp, &yp, &zp);
  return ((int32_t)xp * n + yp) * n + zp;
}

// Occupancy of the 26 neighbours of cell (cx, cy, cz): bit q the half-shell
// cell q, bit 13 + q the mirrored half-shell cell q.  The 26 reads are
// independent and go out together (one round trip to L2, where the mask
// lives), so the list passes then read the runs of the few occupied cells
// only.
__device__ __forceinline__ uint32_t shell_mask(const Dev &d, const Geo &g, int cx, int cy,
                                               int cz) {
  const int n = g.n1d;
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < 13; ++q) {
    int x, y, z;
    shell_cell<false>(n, cx, cy, cz, q, &x, &y, &z);
    m |= (uint32_t)cell_occupied(d, g, x, y, z) << q;
    shell_cell<true>(n, cx, cy, cz, q, &x, &y, &z);
    m |= (uint32_t)cell_occupied(d, g, x, y, z) << (13 + q);
  }
  return m;
}

// The bonds b-1 and b+1 when they share a site with b (Bond::HasNeighbor
// within one mesh, src/bond.cpp:118-121): the pairs the reference culls on
// the rebuild step (src/interaction_manager.cpp:485-497), never listed here.
struct Excl {
  int32_t lo, hi;  // b-1 / b+1, or -1
};

__device__ __forceinline__ Excl excl_of(const Dev &d, const Geo &g, int32_t b) {
  const int32_t f = d.bond_fil[b];
  Excl x;
  x.lo = (b > 0 && d.bond_fil[b - 1] == f) ? b - 1 : -1;
  x.hi = (b + 1 < g.nb && d.bond_fil[b + 1] == f) ? b + 1 : -1;
  return x;
}

// Degree of every bond: the candidates that list it (Cell::MakePairs,
// src/cell.cpp:14-70): the other members of its cell, the members of its 13
// half-shell cells and of its 13 mirrored half-shell cells, less the
// site-sharing neighbours among them.  One thread per sorted position (the
// cell lookups of neighbouring threads then hit the same lines), no
// atomics.  The stencil cells are recomputed per entry rather than held,
// which keeps the register count (and so the occupancy) of this
// latency-bound pass high.  The neighbour occupancy is kept for k_pull.
__device__ __forceinline__ void deg_one(const Dev &d, const Geo &g, int64_t p) {
  const int32_t k = d.skey[p];
  const int32_t b = d.sorted[p];
  const int2 own = cell_run_key(d, g, k);
  const Excl x = excl_of(d, g, b);
  const int32_t klo = x.lo >= 0 ? d.keys[x.lo] : -1;
  const int32_t khi = x.hi >= 0 ? d.keys[x.hi] : -1;
  int32_t deg = (own.y - own.x - 1) - (klo == k) - (khi == k);
  const int n = g.n1d;
  const int cx = k / (n * n), cy = (k / n) % n, cz = k % n;
  const uint32_t m = shell_mask(d, g, cx, cy, cz);
  // the bond's own run, those of its first kRunSlab - 2 occupied stencil
  // cells (in bit order) and the occupancy mask go to the bond's slab (64
  // bytes, one whole line, so the scattered stores waste nothing), where
  // the list pass, which runs in bond order, reads them in coalesced loads
  // without more dependent lookups
  // (the first kRunSlab - 2 lookups are unrolled and independent, so they
  // are in flight together; the slab is stored after them)
  uint32_t mm = m;
  int2 r[kRunSlab - 2];
#pragma unroll
  for (int i = 0; i < kRunSlab - 2; ++i) {
    r[i] = make_int2(0, 0);
    if (mm) {
      const int q = __ffs(mm) - 1;
      mm &= mm - 1;
      int x, y, z;
      if (q < 13)
        shell_cell<false>(n, cx, cy, cz, q, &x, &y, &z);
      else
        shell_cell<true>(n, cx, cy, cz, q - 13, &x, &y, &z);
      const int32_t kc = (x * n + y) * n + z;
      r[i] = cell_run_of(d, g, x, y, z);
      deg += (r[i].y - r[i].x) - (klo == kc) - (khi == kc);
    }
  }
  for (; mm; mm &= mm - 1) {
    const int q = __ffs(mm) - 1;
    int x, y, z;
    if (q < 13)
      shell_cell<false>(n, cx, cy, cz, q, &x, &y, &z);
    else
      shell_cell<true>(n, cx, cy, cz, q - 13, &x, &y, &z);
    const int32_t kc = (x * n + y) * n + z;
    const int2 rr = cell_run_of(d, g, x, y, z);
    deg += (rr.y - rr.x) - (klo == kc) - (khi == kc);
  }
  d.deg[b] = deg;
  int4 *slab = reinterpret_cast<int4 *>(d.nrun + (int64_t)b * kRunSlab);
  static_assert(kRunSlab == 8, "four int4 stores per slab");
  slab[0] = make_int4(own.x, own.y, r[0].x, r[0].y);
  slab[1] = make_int4(r[1].x, r[1].y, r[2].x, r[2].y);
  slab[2] = make_int4(r[3].x, r[3].y, r[4].x, r[4].y);
  slab[3] = make_int4(r[5].x, r[5].y, (int32_t)m, 0);
}

// Grid-stride over the sorted positions: a capped grid keeps the launch
// cheap on the steps that do not rebuild (the guard returns at once).
__device__ __forceinline__ void deg_body(const Dev &d, const Geo &g, int64_t bid, int64_t nblk) {
  const int64_t stride = nblk * blockDim.x;
  for (int64_t p = bid * blockDim.x + threadIdx.x; p < g.nb; p += stride) deg_one(d, g, p);
}

__global__ __launch_bounds__(256) void k_deg(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  deg_body(d, g, blockIdx.x, gridDim.x);
}

// Incidence lists in candidate order, one thread per bond, no atomics and
// no sort.  The list order of CellList::MakePairs (src/cell_list.cpp:
// 127-137, src/cell.cpp:14-70) is: home cells in key order; inside a home
// cell its self pairs (i < j, in member order), then for each of its 13
// half-shell cells in stencil order all (home member, neighbour member)
// pairs, home member outer.  So bond b's candidates come from its own cell
// (as obj2 of the self pairs with earlier members, as obj1 of those with
// later members, then as obj1 of its rows of the 13 half-shell blocks) and
// from each occupied mirrored half-shell cell m (as obj2 of m's block
// toward b's cell, one pair per member of m, in member order), the groups
// taken in home-key order.  That order of the 14 home cells depends only on
// which of b's cell coordinates sit on a periodic face (Geo::home_order,
// built on the host), so no per-thread sort is needed.  Entries {partner,
// role | written << 1 | (b mod 64) << 2}: the owner's place in its 64-bond
// group lets the balanced pair kernel find an entry's bond without a
// search; the slot copies start unwritten (zero force).  The launch also
// checks the capacity, commits the list length and restarts the
// displacement bookkeeping (InteractionManager::ZeroDrTot -> Mesh/Object::
// ZeroDrTot, src/interaction_manager.cpp:316-321, src/mesh.cpp:546-551).
__device__ __forceinline__ void pull_one(const Dev &d, const Geo &g, int32_t b,
                                         int64_t total, int2 *stage, int32_t base) {
  int32_t w = d.inc_start[b];
  const int32_t wend = d.inc_start[b + 1];
  if (w < 0 || wend < w || wend > total) {  // a wrapped count: refuse, never write
    d.ctl->abort = 1;
    d.ctl->overflow = 1;
    return;
  }
  const int32_t k = d.keys[b];
  const Excl x = excl_of(d, g, b);
  const int32_t tag = (b & 63) << 2;
  auto put = [&](int32_t a, int role) {
    if (a == x.lo || a == x.hi) return;
    if (w < wend) {
      if (stage)
        stage[w - base] = make_int2(a, role | tag);
      else
        d.inc[w] = make_int2(a, role | tag);
    }
    ++w;
  };
  const int n = g.n1d;
  const int cx = k / (n * n), cy = (k / n) % n, cz = k % n;
  const int cls = (cx == 0 ? 0 : cx == n - 1 ? 2 : 1) * 6 +
                  (cy == 0 ? 0 : cy == n - 1 ? 2 : 1) * 2 + (cz == 0 ? 0 : 1);
  // the runs k_deg kept: own cell first, then the occupied stencil cells
  // in bit order, and the occupancy mask of the 26 stencil cells; a run
  // beyond the slab is looked up in the cell list
  const int4 *sl = reinterpret_cast<const int4 *>(d.nrun + (int64_t)b * kRunSlab);
  static_assert(kRunSlab == 8, "four int4 loads per slab");
  const int4 s0 = sl[0], s1 = sl[1], s2 = sl[2], s3 = sl[3];
  const int2 own = make_int2(s0.x, s0.y);
  const uint32_t m = (uint32_t)s3.z;
  auto run_of = [&](int q, int32_t key) -> int2 {
    const int idx = 1 + __popc(m & ((1u << q) - 1u));
    switch (idx) {
      case 1: return make_int2(s0.z, s0.w);
      case 2: return make_int2(s1.x, s1.y);
      case 3: return make_int2(s1.z, s1.w);
      case 4: return make_int2(s2.x, s2.y);
      case 5: return make_int2(s2.z, s2.w);
      case 6: return make_int2(s3.x, s3.y);
      default: return d.cell_run[key];
    }
  };
  uint64_t order = g.home_order[cls];
  for (int step = 0; step < 14; ++step, order >>= 4) {
    const int c = (int)(order & 15);
    if (c == 13) {
      // own cell (members in bond order): self pairs, then the rows of the
      // half-shell blocks
      // (a cell of one member holds b alone: no load needed)
      int32_t r = own.x;
      if (own.y - own.x > 1)
        while (r < own.y && d.sorted[r] != b) ++r;
      for (int32_t j = own.x; j < r; ++j) put(d.sorted[j], 1);
      for (int32_t j = r + 1; j < own.y; ++j) put(d.sorted[j], 0);
      for (uint32_t mm = m & 0x1fffu; mm; mm &= mm - 1) {
        const int q = __ffs(mm) - 1;
        const int2 rn = run_of(q, shell_key<false>(n, cx, cy, cz, q));
        for (int32_t j = rn.x; j < rn.y; ++j) put(d.sorted[j], 0);
      }
    } else if ((m >> (13 + c)) & 1u) {
      // mirrored half-shell cell c: b is obj2 of each of its members
      const int2 rn = run_of(13 + c, shell_key<true>(n, cx, cy, cz, c));
      for (int32_t j = rn.x; j < rn.y; ++j) put(d.sorted[j], 1);
    }
  }
  if (w != wend) {  // degree and list disagree: never expected; fail the step
    d.ctl->abort = 1;
    d.ctl->err = SC_ERR_STATE;
  }
}

constexpr int kPullStage = 512;

__device__ __forceinline__ void pull_body(const Dev &d, const Geo &g, int64_t bid, int64_t nblk) {
  const int64_t i0 = bid * blockDim.x + threadIdx.x;
  const int64_t stride = nblk * blockDim.x;
  const int64_t total = d.inc_start[g.nb];  // two incidences per candidate
  const bool fits = total >= 0 && total <= 2 * g.cap;
  if (i0 == 0) {
    d.ctl->nc_need = total >= 0 ? total / 2 : 2 * g.cap + 1;
    if (fits) {
      d.ctl->nc = total / 2;
      d.ctl->n_rebuilds += 1;
    } else {
      d.ctl->abort = 1;
      d.ctl->overflow = 1;
    }
  }
  if (!fits) return;
  for (int64_t i = i0; i < g.ns; i += stride) {
    for (int q = 0; q < 3; ++q) d.dr0[q * g.ns + i] = d.r[q * g.ns + i];
    d.acc[i] = 0.0;
  }
  // each wavefront takes 64 consecutive bonds, whose lists are one
  // contiguous range: written to LDS entry by entry, then copied out in
  // coalesced stores (ranges longer than the stage are written directly)
  __shared__ int2 stage[4][kPullStage];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t ngrp = (g.nb + 255) / 256;
  for (int64_t grp = bid; grp < ngrp; grp += nblk) {
    const int64_t b0 = grp * 256 + wv * 64, b = b0 + lane;
    int32_t base = 0, cnt = 0;
    if (b0 < g.nb) {
      base = d.inc_start[b0];
      cnt = d.inc_start[b0 + 64 < g.nb ? b0 + 64 : g.nb] - base;
    }
    const bool staged = cnt >= 0 && cnt <= kPullStage;
    if (b < g.nb) pull_one(d, g, (int32_t)b, total, staged ? stage[wv] : nullptr, base);
    __syncthreads();
    if (staged)
      for (int i = lane; i < cnt; i += 64) d.inc[base + i] = stage[wv][i];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_pull(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  pull_body(d, g, blockIdx.x, gridDim.x);
}

static inline unsigned nblocks(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  return (unsigned)(b > 0 ? b : 1);
}

static inline unsigned sgrid(int64_t n) {
  const unsigned b = nblocks(n, 256);
  return b < (unsigned)kStreamBlocks ? b : (unsigned)kStreamBlocks;
}

// keys (+ column populations) -> column scan (2) -> column scatter ->
// column sort (+ runs, occupancy) -> degrees -> degree scan (2) ->
// incidence lists (+ capacity check, displacement restart): 9 launches,
// none over the n1d^3 cells.  Every launch's grid is capped (the kernels
// loop over their range), so on the odd steps that do not rebuild each
// launch only reads the guard.  The two per-bond list passes are
// latency-bound and get 8 waves per SIMD.
constexpr unsigned kListBlocks = 2048;

void launch_rebuild(const Dev &d, const Geo &g, hipStream_t st) {
  const unsigned gs = sgrid(g.nb);
  const int64_t ncol = (int64_t)g.n1d * g.n1d;
  hipLaunchKernelGGL(k_keys, dim3(gs), dim3(kScanBlock), 0, st, d, g);
  launch_scan(d.col_count, d.col_start, ncol, d.scan_tmp, d.ctl, kGuardRebuild, st);
  hipLaunchKernelGGL(k_scatter, dim3(gs), dim3(kScanBlock), 0, st, d, g, (int)kGuardRebuild);
  hipLaunchKernelGGL(k_colsort, dim3(colsort_grid(ncol)), dim3(64 * kColWaves), 0, st, d, g,
                     (int)kGuardRebuild);
  const unsigned gl = std::min(nblocks(g.nb, 256), kListBlocks);
  hipLaunchKernelGGL(k_deg, dim3(gl), dim3(256), 0, st, d, g);
  launch_scan(d.deg, d.inc_start, g.nb, d.scan_tmp, d.ctl, kGuardRebuild, st);
  hipLaunchKernelGGL(k_pull, dim3(std::min(nblocks(g.ns > g.nb ? g.ns : g.nb, 256), kListBlocks)),
                     dim3(256), 0, st, d, g);
}

// Ensemble batches: the launches of launch_rebuild, each over all members
// (odd-step pointers: a rebuild only ever starts an odd step); a member
// that does not rebuild returns at its guard.  Per member, the same grid as
// its own launch.
__global__ void k_group_keys(const Member *__restrict__ ms, GroupGrid gg) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  if (guard_off(M.d[0].ctl, kGuardRebuild)) return;
  keys_body(M.d[0], M.g, blockIdx.x - b0, nblk);
}

__global__ void k_group_scatter(const Member *__restrict__ ms, GroupGrid gg) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  if (guard_off(M.d[0].ctl, kGuardRebuild)) return;
  scatter_body(M.d[0], M.g, blockIdx.x - b0, nblk);
}

__global__ __launch_bounds__(64) void k_group_colsort(const Member *__restrict__ ms,
                                                      GroupGrid gg) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  if (guard_off(M.d[0].ctl, kGuardRebuild)) return;
  colsort_body<1>(M.d[0], M.g, blockIdx.x - b0, nblk);
}

__global__ __launch_bounds__(256) void k_group_deg(const Member *__restrict__ ms, GroupGrid gg) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  if (guard_off(M.d[0].ctl, kGuardRebuild)) return;
  deg_body(M.d[0], M.g, blockIdx.x - b0, nblk);
}

__global__ __launch_bounds__(256) void k_group_pull(const Member *__restrict__ ms, GroupGrid gg) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  if (guard_off(M.d[0].ctl, kGuardRebuild)) return;
  pull_body(M.d[0], M.g, blockIdx.x - b0, nblk);
}

// which 0: the column populations into column starts; 1: the degrees into
// the incidence starts
__global__ __launch_bounds__(kScanBlock) void k_group_scan_reduce(const Member *__restrict__ ms,
                                                                  GroupGrid gg, int which) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  const Dev &d = M.d[0];
  if (guard_off(d.ctl, kGuardRebuild)) return;
  const int64_t n = which ? M.g.nb : (int64_t)M.g.n1d * M.g.n1d;
  scan_reduce_body(which ? d.deg : d.col_count, n, d.scan_tmp, blockIdx.x - b0, nblk);
}

__global__ __launch_bounds__(kScanBlock) void k_group_scan_down(const Member *__restrict__ ms,
                                                                GroupGrid gg, int which) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  const Dev &d = M.d[0];
  if (guard_off(d.ctl, kGuardRebuild)) return;
  const int64_t n = which ? M.g.nb : (int64_t)M.g.n1d * M.g.n1d;
  scan_down_body(which ? d.deg : d.col_count, n, d.scan_tmp, which ? d.inc_start : d.col_start,
                 blockIdx.x - b0, nblk);
}

namespace {
// offsets of a grouped launch from each member's block count; the total
template <typename F>
int32_t group_grid(GroupGrid &gg, const Member *hm, int n, F blocks) {
  int32_t t = 0;
  for (int m = 0; m < n; ++m) {
    gg.off[m] = t;
    t += (int32_t)blocks(hm[m].g);
  }
  for (int m = n; m <= kMaxGroup; ++m) gg.off[m] = t;
  return t;
}

unsigned scan_blocks(int64_t n) {
  int64_t nblk = (n + kScanTile - 1) / kScanTile;
  if (nblk == 0) nblk = 1;
  return (unsigned)(nblk > kStreamBlocks ? kStreamBlocks : nblk);
}
}  // namespace

void launch_group_rebuild(const Member *dm, const Member *hm, int n, hipStream_t st) {
  if (n <= 0) return;
  GroupGrid gk, gcs, gsrt, gdeg, gds, gpl;
  const int32_t tk = group_grid(gk, hm, n, [](const Geo &g) { return sgrid(g.nb); });
  const int32_t tcs = group_grid(gcs, hm, n, [](const Geo &g) {
    return scan_blocks((int64_t)g.n1d * g.n1d);
  });
  const int32_t tsrt = group_grid(gsrt, hm, n, [](const Geo &g) {
    return (unsigned)std::min<int64_t>((int64_t)g.n1d * g.n1d, 8192);
  });
  const int32_t tdeg = group_grid(gdeg, hm, n, [](const Geo &g) {
    return std::min(nblocks(g.nb, 256), kListBlocks);
  });
  const int32_t tds = group_grid(gds, hm, n, [](const Geo &g) { return scan_blocks(g.nb); });
  const int32_t tpl = group_grid(gpl, hm, n, [](const Geo &g) {
    return std::min(nblocks(g.ns > g.nb ? g.ns : g.nb, 256), kListBlocks);
  });
  hipLaunchKernelGGL(k_group_keys, dim3(tk), dim3(kScanBlock), 0, st, dm, gk);
  hipLaunchKernelGGL(k_group_scan_reduce, dim3(tcs), dim3(kScanBlock), 0, st, dm, gcs, 0);
  hipLaunchKernelGGL(k_group_scan_down, dim3(tcs), dim3(kScanBlock), 0, st, dm, gcs, 0);
  hipLaunchKernelGGL(k_group_scatter, dim3(tk), dim3(kScanBlock), 0, st, dm, gk);
  hipLaunchKernelGGL(k_group_colsort, dim3(tsrt), dim3(64), 0, st, dm, gsrt);
  hipLaunchKernelGGL(k_group_deg, dim3(tdeg), dim3(256), 0, st, dm, gdeg);
  hipLaunchKernelGGL(k_group_scan_reduce, dim3(tds), dim3(kScanBlock), 0, st, dm, gds, 1);
  hipLaunchKernelGGL(k_group_scan_down, dim3(tds), dim3(kScanBlock), 0, st, dm, gds, 1);
  hipLaunchKernelGGL(k_group_pull, dim3(tpl), dim3(256), 0, st, dm, gpl);
}

void launch_group_commit(const Member *dm, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_group_commit, dim3(n), dim3(64), 0, st, dm);
}

void launch_group_ctl(const Member *dm, int n, Ctl *out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_group_ctl, dim3(1), dim3(64), 0, st, dm, n, out);
}

// The cell list of the last rebuild (sorted, runs, occupancy) from the
// stored keys; after a checkpoint restore the crosslink partner search
// needs the list the saved state was stepping with.
void launch_cells_from_keys(const Dev &d, const Geo &g, hipStream_t st) {
  const int64_t ncol = (int64_t)g.n1d * g.n1d;
  hipLaunchKernelGGL(k_key_count, dim3(sgrid(g.nb)), dim3(256), 0, st, d, g);
  launch_scan(d.col_count, d.col_start, ncol, d.scan_tmp, nullptr, 0, st);
  hipLaunchKernelGGL(k_scatter, dim3(sgrid(g.nb)), dim3(256), 0, st, d, g, 0);
  hipLaunchKernelGGL(k_colsort, dim3(colsort_grid(ncol)), dim3(64 * kColWaves), 0, st, d, g, 0);
}

// Replays the displacement bookkeeping of the last full step after the host
// replaced the site positions: acc = acc_prev + 2 |r - r0|^2 with the new
// positions, the order Object::UpdateDrTot adds in (src/object.cpp:282-287).
__global__ void k_redo_dr(Dev d, Geo g) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0;
  if (i < g.ns) {
    a = d.acc_prev[i];
    for (int rep = 0; rep < 2; ++rep)
      for (int k = 0; k < 3; ++k) {
        double dd = d.r[k * g.ns + i] - d.dr0[k * g.ns + i];
        a += dd * dd;
      }
    d.acc[i] = a;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double y = __shfl_xor(a, o, 64);
    a = y > a ? y : a;
  }
  if ((threadIdx.x & 63) == 0)
    atomicMax(&d.dr_slots[(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) &
                          (kDrSlots - 1)],
              (unsigned long long)__double_as_longlong(a));
}

void launch_redo_dr(const Dev &d, const Geo &g, hipStream_t st) {
  hipLaunchKernelGGL(k_redo_dr, dim3(nblocks(g.ns, 256)), dim3(256), 0, st, d, g);
}

// Plain streaming copy, the calibration kernel for HBM byte counters.
__global__ void k_copy_probe(const double4 *__restrict__ src, double4 *__restrict__ dst,
                             int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[i];
}

void launch_copy_probe(const void *src, void *dst, int64_t n32, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, st, (const double4 *)src,
                     (double4 *)dst, n32);
}

void launch_decide(const Dev &d, const Geo &g, hipStream_t st) {
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, st, d.ctl, d.dr_slots, g.thr);
}

void launch_commit(const Dev &d, const Geo &g, hipStream_t st) {
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(64), 0, st, d.ctl, d.dr_slots, g.thr);
}

void launch_zero_i32(int32_t *p, int64_t n, hipStream_t st) {
  unsigned gb = nblocks(n, 256);
  if (gb > 4096) gb = 4096;
  hipLaunchKernelGGL(k_zero_i32, dim3(gb), dim3(256), 0, st, p, n);
}

__global__ void k_bond_init(Dev d, Geo g, int prec) {
  int f = blockIdx.x * blockDim.y + threadIdx.y;
  if (f >= g.nf) return;
  const int64_t s0 = d.site_off[f], s1 = d.site_off[f + 1];
  const int64_t b0 = s0 - f;
  for (int64_t i = threadIdx.x; i < s1 - s0 - 1; i += blockDim.x) {
    double r1[3], r2[3];
    for (int k = 0; k < 3; ++k) {
      r1[k] = d.r[k * g.ns + s0 + i];
      r2[k] = d.r[k * g.ns + s0 + i + 1];
    }
    bond_reinit(d, g, b0 + i, r1, r2, prec);
  }
}

void launch_bond_init(const Dev &d, const Geo &g, int precision,
                      hipStream_t st) {
  dim3 bs(64, 4);
  hipLaunchKernelGGL(k_bond_init, dim3(nblocks(g.nf, 4)), bs, 0, st, d, g,
                     precision);
}

}  // namespace sc
