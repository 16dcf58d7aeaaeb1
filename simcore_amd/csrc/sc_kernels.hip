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
//   [rebuild: keys, scan, scatter, count (+ run order, degrees), emit
//    (+ offsets, cap), scan, incidence fill, incidence sort, zero dr]
//   -> pair+scatter -> integrate (one launch per length bin)
//   -> commit (and decide the next step)
#include <hip/hip_runtime.h>

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
__global__ void k_commit(Ctl *c, unsigned long long *slots, double thr) {
  static_assert(kDrSlots == 4 * 64, "four slots per lane");
  const int lane = threadIdx.x;
  unsigned long long v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = slots[lane + 64 * i];
  const int32_t abort = c->abort;
  const int64_t cur = c->cur_step;
  const unsigned long long drb = c->dr_bits;
  if (blockIdx.x != 0 || abort) return;
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

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(
    const int32_t *in, int64_t n, int32_t *bsum, const Ctl *ctl, int guard) {
  if (guard_off(ctl, guard)) return;
  __shared__ int wsum[kScanBlock / 64];
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
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
// With `occ` set, the pass also writes one bit per entry, set when the
// entry (a cell's population) is non-zero: the cell-occupancy mask.
__global__ __launch_bounds__(kScanBlock) void k_scan_down(const int32_t *in,
                                                          int64_t n,
                                                          const int32_t *bsum,
                                                          int32_t *out,
                                                          const Ctl *ctl,
                                                          int guard, uint8_t *occ) {
  if (guard_off(ctl, guard)) return;
  __shared__ int tile[kScanTile];
  __shared__ int wsum[kScanBlock / 64];
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int part = 0;
    for (int64_t i = threadIdx.x; i < t; i += kScanBlock) part += bsum[i];
    int carry;
    block_excl_scan256(part, &carry, wsum);
    const int64_t base = t * kScanTile;
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
      tile[j] = (base + j < n) ? in[base + j] : 0;
    __syncthreads();
    if (occ && base + threadIdx.x * kScanItems < n) {
      static_assert(kScanItems == 8, "one mask byte per thread");
      unsigned bits = 0;
#pragma unroll
      for (int j = 0; j < kScanItems; ++j)
        bits |= (tile[threadIdx.x * kScanItems + j] != 0 ? 1u : 0u) << j;
      occ[(base >> 3) + threadIdx.x] = (uint8_t)bits;
    }
    int tot;
    tile_scan(tile, wsum, carry, &tot);
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
      if (base + j < n) out[base + j] = tile[j];
    if (t == ntiles - 1 && threadIdx.x == 0) out[n] = carry + tot;
    __syncthreads();
  }
}

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *tmp,
                 const Ctl *ctl, int guard, hipStream_t st, uint8_t *occ) {
  int64_t nblk = (n + kScanTile - 1) / kScanTile;
  if (nblk == 0) nblk = 1;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblk), dim3(kScanBlock), 0,
                     st, in, n, tmp, ctl, guard);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nblk), dim3(kScanBlock), 0,
                     st, in, n, tmp, out, ctl, guard, occ);
}

// ---------------------------------------------------------------------------
// Rebuild stage
// ---------------------------------------------------------------------------

// CellList::FindCellCoords over every bond (src/cell_list.cpp:139-157) and
// the per-cell population count.  Streaming passes loop over their range,
// so any grid size serves.
__global__ void k_keys(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < g.nb; b += stride) {
    int32_t k;
    if (g.compact) {
      k = rec_key_compact(d.br, g, b);
    } else {
      const double *rec = d.br + 8 * b;
      k = cell_key(g.n1d, rec[0], rec[1], rec[2]);
    }
    d.keys[b] = k;
    atomicAdd(&d.cell_count[k], 1);
  }
}

// Population count from the stored keys (cell list restore).
__global__ void k_key_count(Dev d, Geo g) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < g.nb; b += stride)
    atomicAdd(&d.cell_count[d.keys[b]], 1);
}

// Counting-sort scatter (order inside a cell arbitrary; k_deg orders the
// runs); returns cell_count to all-zero for the next build.
__global__ void k_scatter(Dev d, Geo g, int guard) {
  if (guard_off(d.ctl, guard)) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < g.nb; b += stride) {
    const int32_t k = d.keys[b];
    const int32_t pos = d.cell_start[k] + atomicSub(&d.cell_count[k], 1) - 1;
    d.sorted[pos] = (int32_t)b;
    d.skey[pos] = k;
  }
}

// q-th cell of the half-shell stencil of (x, y, z) in the order of
// CellList::AssignCellNeighbors (src/cell_list.cpp:206-262): (x+1, y, z);
// (x-1..x+1, y+1, z); then (x-1..x+1, y-1..y+1, z+1), x fastest.
__device__ __forceinline__ int32_t shell_key(int n, int x, int y, int z, int q) {
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
  int xp = x + dx, yp = y + dy, zp = z + dz;
  xp = xp < 0 ? n - 1 : (xp == n ? 0 : xp);
  yp = yp < 0 ? n - 1 : (yp == n ? 0 : yp);
  zp = zp == n ? 0 : zp;
  return ((int32_t)xp * n + yp) * n + zp;
}

// q-th cell of the mirrored half shell: the 13 cells whose half shell holds
// (x, y, z), i.e. the offsets of shell_key negated (wrapped).  With at
// least 3 cells per side the 13 + 13 + 1 cells are distinct.
__device__ __forceinline__ int32_t back_shell_key(int n, int x, int y, int z, int q) {
  int dx, dy, dz;
  if (q == 0) {
    dx = -1;
    dy = 0;
    dz = 0;
  } else if (q < 4) {
    dx = 2 - q;
    dy = -1;
    dz = 0;
  } else {
    dx = 1 - (q - 4) % 3;
    dy = 1 - (q - 4) / 3;
    dz = -1;
  }
  int xp = x + dx, yp = y + dy, zp = z + dz;
  xp = xp < 0 ? n - 1 : (xp == n ? 0 : xp);
  yp = yp < 0 ? n - 1 : (yp == n ? 0 : yp);
  zp = zp < 0 ? n - 1 : zp;
  return ((int32_t)xp * n + yp) * n + zp;
}

// Run [start, end) of cell c in sorted / sorted_o, empty for an unoccupied
// cell.  Most cells are empty; they are recognised from the occupancy mask,
// a small array that stays in L2, so only occupied ones read the large
// cell_start array.
__device__ __forceinline__ void cell_run(const Dev &d, int32_t c, int32_t *s, int32_t *e) {
  *s = 0;
  *e = 0;
  if ((d.cell_occ[c >> 3] >> (c & 7)) & 1u) {
    *s = d.cell_start[c];
    *e = d.cell_start[c + 1];
  }
}

// The 13 half-shell cells (fw) or mirrored half-shell cells (bw) of k with
// their runs, loaded up front so that the lookups are in flight together.
struct Shell {
  int32_t key[13], o0[13], o1[13];
};

template <bool BACK>
__device__ __forceinline__ void load_shell(const Dev &d, const Geo &g, int32_t k, Shell *sh) {
  const int n = g.n1d;
  const int x = k / (n * n), y = (k / n) % n, z = k % n;
#pragma unroll
  for (int q = 0; q < 13; ++q)
    sh->key[q] = BACK ? back_shell_key(n, x, y, z, q) : shell_key(n, x, y, z, q);
#pragma unroll
  for (int q = 0; q < 13; ++q) cell_run(d, sh->key[q], &sh->o0[q], &sh->o1[q]);
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
// site-sharing neighbours among them.  One thread per sorted position, no
// atomics; it also places its bond in its cell's run of sorted_o by its
// rank in bond order (Cell::AddObj appends in interactor order,
// src/cell.cpp:5-8), the order the pair lists are written in.
__global__ __launch_bounds__(256) void k_deg(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.nb) return;
  const int32_t k = d.skey[p];
  const int32_t b = d.sorted[p];
  const int32_t s0 = d.cell_start[k], e = d.cell_start[k + 1];
  const Excl x = excl_of(d, g, b);
  const int32_t klo = x.lo >= 0 ? d.keys[x.lo] : -1;
  const int32_t khi = x.hi >= 0 ? d.keys[x.hi] : -1;
  Shell fw, bw;
  load_shell<false>(d, g, k, &fw);
  load_shell<true>(d, g, k, &bw);
  int32_t rank = 0;
  for (int32_t i = s0; i < e; ++i) rank += d.sorted[i] < b;
  d.sorted_o[s0 + rank] = b;
  int32_t deg = (e - s0 - 1) - (klo == k) - (khi == k);
#pragma unroll
  for (int q = 0; q < 13; ++q) {
    deg += (fw.o1[q] - fw.o0[q]) - (klo == fw.key[q]) - (khi == fw.key[q]);
    deg += (bw.o1[q] - bw.o0[q]) - (klo == bw.key[q]) - (khi == bw.key[q]);
  }
  d.deg[b] = deg;
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
// taken in home-key order.  Entries {partner, role}; the slot copies start
// unwritten (zero force).  The launch also checks the capacity, commits the
// list length and restarts the displacement bookkeeping
// (InteractionManager::ZeroDrTot -> Mesh/Object::ZeroDrTot,
// src/interaction_manager.cpp:316-321, src/mesh.cpp:546-551).
__global__ __launch_bounds__(256) void k_pull(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = d.inc_start[g.nb];  // two incidences per candidate
  const bool fits = total >= 0 && total <= 2 * g.cap;
  if (i == 0) {
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
  if (i < g.ns) {
    for (int q = 0; q < 3; ++q) d.dr0[q * g.ns + i] = d.r[q * g.ns + i];
    d.acc[i] = 0.0;
  }
  if (i >= g.nb) return;
  const int32_t b = (int32_t)i;
  int32_t w = d.inc_start[b];
  const int32_t wend = d.inc_start[b + 1];
  if (w < 0 || wend < w || wend > total) {  // a wrapped count: refuse, never write
    d.ctl->abort = 1;
    d.ctl->overflow = 1;
    return;
  }
  const int32_t k = d.keys[b];
  const Excl x = excl_of(d, g, b);
  int32_t s0, e;
  cell_run(d, k, &s0, &e);
  Shell fw, bw;
  load_shell<false>(d, g, k, &fw);
  load_shell<true>(d, g, k, &bw);
  auto put = [&](int32_t a, int role) {
    if (a == x.lo || a == x.hi) return;
    if (w < wend) d.inc[w] = make_int2(a, role);
    ++w;
  };
  // home cells in key order: b's own cell k and the occupied mirrored
  // cells, by repeated selection of the next larger key (13 is small)
  int32_t last = -1;
  for (int step = 0; step < 14; ++step) {
    int32_t next = 0x7fffffff;
    int which = -2;  // -1: own cell, q: mirrored cell q
    if (k > last) {
      next = k;
      which = -1;
    }
#pragma unroll
    for (int q = 0; q < 13; ++q)
      if (bw.o1[q] > bw.o0[q] && bw.key[q] > last && bw.key[q] < next) {
        next = bw.key[q];
        which = q;
      }
    if (which == -2) break;
    last = next;
    if (which == -1) {
      int32_t r = s0;
      while (r < e && d.sorted_o[r] != b) ++r;
      for (int32_t j = s0; j < r; ++j) put(d.sorted_o[j], 1);
      for (int32_t j = r + 1; j < e; ++j) put(d.sorted_o[j], 0);
#pragma unroll
      for (int q = 0; q < 13; ++q)
        for (int32_t j = fw.o0[q]; j < fw.o1[q]; ++j) put(d.sorted_o[j], 0);
    } else {
#pragma unroll
      for (int q = 0; q < 13; ++q)
        if (q == which)
          for (int32_t j = bw.o0[q]; j < bw.o1[q]; ++j) put(d.sorted_o[j], 1);
    }
  }
  if (w != wend) {  // degree and list disagree: never expected; fail the step
    d.ctl->abort = 1;
    d.ctl->err = SC_ERR_STATE;
  }
}

static inline unsigned nblocks(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  return (unsigned)(b > 0 ? b : 1);
}

static inline unsigned sgrid(int64_t n) {
  const unsigned b = nblocks(n, 256);
  return b < (unsigned)kStreamBlocks ? b : (unsigned)kStreamBlocks;
}

// keys (+ cell populations) -> cell scan (2) -> scatter -> degrees (+ run
// order) -> degree scan (2) -> incidence lists (+ capacity check,
// displacement restart): 8 launches.
void launch_rebuild(const Dev &d, const Geo &g, hipStream_t st) {
  const unsigned gb = nblocks(g.nb, kScanBlock);
  hipLaunchKernelGGL(k_keys, dim3(gb), dim3(kScanBlock), 0, st, d, g);
  launch_scan(d.cell_count, d.cell_start, g.n_cells, d.scan_tmp, d.ctl,
              kGuardRebuild, st, d.cell_occ);
  hipLaunchKernelGGL(k_scatter, dim3(gb), dim3(kScanBlock), 0, st, d, g, (int)kGuardRebuild);
  hipLaunchKernelGGL(k_deg, dim3(nblocks(g.nb, 256)), dim3(256), 0, st, d, g);
  launch_scan(d.deg, d.inc_start, g.nb, d.scan_tmp, d.ctl, kGuardRebuild, st);
  hipLaunchKernelGGL(k_pull, dim3(nblocks(g.ns > g.nb ? g.ns : g.nb, 256)), dim3(256), 0, st,
                     d, g);
}

// Cell membership (cell_start, sorted) of the last rebuild from the stored
// keys; after a checkpoint restore the crosslink partner search needs the
// list the saved state was stepping with.
void launch_cells_from_keys(const Dev &d, const Geo &g, hipStream_t st) {
  hipLaunchKernelGGL(k_key_count, dim3(sgrid(g.nb)), dim3(256), 0, st, d, g);
  launch_scan(d.cell_count, d.cell_start, g.n_cells, d.scan_tmp, nullptr, 0, st, d.cell_occ);
  hipLaunchKernelGGL(k_scatter, dim3(sgrid(g.nb)), dim3(256), 0, st, d, g, 0);
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
