// sc_kernels.hip -- gfx950 kernels of the filament BD step.
//
// Built with -ffp-contract=off: every FP64 multiply and add rounds on its
// own, in the order the reference's source states it, so the FP64 path
// reproduces the CPU path value-for-value.  Division and sqrt on FP64 are
// IEEE correctly rounded on gfx950 (v_div_scale/fmas/fixup sequence and
// the refined v_sqrt_f64 expansion).
//
// One step, in stream order (all guarded by the device control block):
//   decide -> [rebuild: keys, scan, scatter, run sort, count, scan, cap,
//              emit, scan, incidence fill, incidence sort, zero dr]
//          -> pair -> reduce -> integrate (one launch per length bin)
//          -> commit
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

// ---------------------------------------------------------------------------
// Control kernels
// ---------------------------------------------------------------------------

// Replaces InteractionManager::CheckUpdateInteractions / GetDrMax
// (src/interaction_manager.cpp:284-312): the accumulator of a full step is
// compared with 0.25*cell_len^2 at the next (odd) step; the mesh skips the
// update after a midstep (src/mesh.cpp:533-536).
__global__ void k_decide(Ctl *c, double thr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->abort) return;
  int64_t k = c->i_step + 1;
  c->cur_step = k;
  c->midstep = (k & 1) ? 1 : 0;
  int rb = 0;
  if ((k & 1) && k >= 3) {
    double drm = __longlong_as_double((long long)c->dr_bits);
    if (drm > 10000) {  // include/simcore/species.hpp:114-118
      c->abort = 1;
      c->err = SC_ERR_UNPHYSICAL;
      return;
    }
    if (drm > thr) rb = 1;
  }
  if (!(k & 1)) c->dr_bits = 0ull;  // the full step about to run re-accumulates
  c->rebuild = rb;
}

__global__ void k_commit(Ctl *c) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->abort) return;
  c->i_step = c->cur_step;
}

__global__ void k_zero_i32(int32_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = 0;
}

__device__ __forceinline__ bool guard_off(const Ctl *c, int guard) {
  if (!c) return false;
  if (c->abort) return true;
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
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + (int64_t)j * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  int tot;
  block_excl_scan256(s, &tot, wsum);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
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

__global__ __launch_bounds__(kScanBlock) void k_scan_mid(int32_t *bsum,
                                                         int64_t nblk,
                                                         int32_t *out,
                                                         int64_t n,
                                                         const Ctl *ctl,
                                                         int guard) {
  if (guard_off(ctl, guard)) return;
  __shared__ int tile[kScanTile];
  __shared__ int wsum[kScanBlock / 64];
  int carry = 0;
  for (int64_t b0 = 0; b0 < nblk; b0 += kScanTile) {
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
      tile[j] = (b0 + j < nblk) ? bsum[b0 + j] : 0;
    __syncthreads();
    int tot;
    tile_scan(tile, wsum, carry, &tot);
    for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
      if (b0 + j < nblk) bsum[b0 + j] = tile[j];
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const int32_t *in,
                                                          int64_t n,
                                                          const int32_t *bsum,
                                                          int32_t *out,
                                                          const Ctl *ctl,
                                                          int guard) {
  if (guard_off(ctl, guard)) return;
  __shared__ int tile[kScanTile];
  __shared__ int wsum[kScanBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
    tile[j] = (base + j < n) ? in[base + j] : 0;
  __syncthreads();
  int tot;
  tile_scan(tile, wsum, bsum[blockIdx.x], &tot);
  for (int j = threadIdx.x; j < kScanTile; j += kScanBlock)
    if (base + j < n) out[base + j] = tile[j];
}

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *tmp,
                 const Ctl *ctl, int guard, hipStream_t st) {
  int64_t nblk = (n + kScanTile - 1) / kScanTile;
  if (nblk == 0) nblk = 1;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblk), dim3(kScanBlock), 0,
                     st, in, n, tmp, ctl, guard);
  hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(kScanBlock), 0, st, tmp, nblk,
                     out, n, ctl, guard);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nblk), dim3(kScanBlock), 0,
                     st, in, n, tmp, out, ctl, guard);
}

// ---------------------------------------------------------------------------
// Rebuild stage
// ---------------------------------------------------------------------------

// CellList::FindCellCoords over every bond (src/cell_list.cpp:139-157) and
// the per-cell population count.
__global__ void k_keys(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nb) return;
  int32_t k = cell_key(g.n1d, d.spos[b], d.spos[g.nb + b], d.spos[2 * g.nb + b]);
  d.keys[b] = k;
  atomicAdd(&d.cell_count[k], 1);
}

// Counting-sort scatter; returns cell_count to all-zero for the next build.
__global__ void k_scatter(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nb) return;
  int32_t k = d.keys[b];
  int32_t pos = d.cell_start[k] + atomicSub(&d.cell_count[k], 1) - 1;
  d.sorted[pos] = (int32_t)b;
  d.skey[pos] = k;
}

// Cell::AddObj appends in interactor (bond index) order; the scatter above
// is unordered, so each cell's run is sorted by bond index.
__global__ void k_run_sort(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.nb) return;
  int32_t k = d.skey[p];
  if (d.cell_start[k] != p) return;
  int32_t e = d.cell_start[k + 1];
  for (int32_t i = (int32_t)p + 1; i < e; ++i) {
    int32_t v = d.sorted[i];
    int32_t j = i - 1;
    while (j >= p && d.sorted[j] > v) {
      d.sorted[j + 1] = d.sorted[j];
      --j;
    }
    d.sorted[j + 1] = v;
  }
}

__device__ __forceinline__ bool excluded(const int32_t *bond_fil, int32_t a,
                                         int32_t b) {
  // Bond::HasNeighbor within one mesh (src/bond.cpp:118-121,
  // src/interaction_manager.cpp:352-367)
  int32_t df = a - b;
  return (df == 1 || df == -1) && bond_fil[a] == bond_fil[b];
}

// Candidate count of each occupied home cell (Cell::MakePairs,
// src/cell.cpp:14-70), excluded pairs already removed (the reference culls
// them on the rebuild step, src/interaction_manager.cpp:485-497).
__global__ void k_count(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.nb) return;
  int32_t k = d.skey[p];
  int32_t cnt = 0;
  if (d.cell_start[k] == p) {
    int32_t e = d.cell_start[k + 1];
    for (int32_t i = (int32_t)p; i < e; ++i)
      for (int32_t j = i + 1; j < e; ++j)
        cnt += !excluded(d.bond_fil, d.sorted[i], d.sorted[j]);
    int32_t nbr[13];
    half_shell(g.n1d, k, nbr);
    for (int q = 0; q < 13; ++q) {
      int32_t o0 = d.cell_start[nbr[q]], o1 = d.cell_start[nbr[q] + 1];
      for (int32_t i = (int32_t)p; i < e; ++i) {
        int32_t bi = d.sorted[i];
        for (int32_t j = o0; j < o1; ++j) cnt += !excluded(d.bond_fil, bi, d.sorted[j]);
      }
    }
  }
  d.ccount[p] = cnt;
}

__global__ void k_check_cap(Dev d, Geo g) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t need = d.coff[g.nb];
  d.ctl->nc_need = need;
  if (need > g.cap) {
    d.ctl->abort = 1;
    d.ctl->overflow = 1;
  } else {
    d.ctl->nc = need;
  }
}

// Writes the candidate list in MakePairs order: home cells ascending; in
// each, the self pairs (i<j) then, for each of the 13 half-shell
// neighbours in AssignCellNeighbors order, all home x neighbour pairs.
__global__ void k_emit(Dev d, Geo g, int prec) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.nb) return;
  int32_t k = d.skey[p];
  if (d.cell_start[k] != p) return;
  int32_t e = d.cell_start[k + 1];
  int32_t w = d.coff[p];
  auto put = [&](int32_t b1, int32_t b2) {
    d.pairs[w] = make_int2(b1, b2);
    if (prec == SC_PREC_FP64) {
      d.slot64[w] = make_double4(0.0, 0.0, 0.0, 0.0);
    } else {
      d.slot32[w] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    atomicAdd(&d.deg[b1], 1);
    atomicAdd(&d.deg[b2], 1);
    ++w;
  };
  for (int32_t i = (int32_t)p; i < e; ++i)
    for (int32_t j = i + 1; j < e; ++j) {
      int32_t bi = d.sorted[i], bj = d.sorted[j];
      if (!excluded(d.bond_fil, bi, bj)) put(bi, bj);
    }
  int32_t nbr[13];
  half_shell(g.n1d, k, nbr);
  for (int q = 0; q < 13; ++q) {
    int32_t o0 = d.cell_start[nbr[q]], o1 = d.cell_start[nbr[q] + 1];
    for (int32_t i = (int32_t)p; i < e; ++i) {
      int32_t bi = d.sorted[i];
      for (int32_t j = o0; j < o1; ++j) {
        int32_t bj = d.sorted[j];
        if (!excluded(d.bond_fil, bi, bj)) put(bi, bj);
      }
    }
  }
}

// Incidence CSR: every candidate appears under obj1 (role 0) and obj2
// (role 1); entries are (c << 1 | role).
__global__ void k_inc_fill(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t nc = d.ctl->nc;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += st) {
    int2 pr = d.pairs[c];
    int32_t p1 = d.inc_start[pr.x] + atomicSub(&d.deg[pr.x], 1) - 1;
    int32_t p2 = d.inc_start[pr.y] + atomicSub(&d.deg[pr.y], 1) - 1;
    d.inc[p1] = (int32_t)(c << 1);
    d.inc[p2] = (int32_t)((c << 1) | 1);
  }
}

// Per-bond incidence lists in candidate order, so the reduction adds the
// contributions in the order ApplyPairInteractions does.
__global__ void k_inc_sort(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nb) return;
  int32_t s = d.inc_start[b], e = d.inc_start[b + 1];
  for (int32_t i = s + 1; i < e; ++i) {
    int32_t v = d.inc[i];
    int32_t j = i - 1;
    while (j >= s && d.inc[j] > v) {
      d.inc[j + 1] = d.inc[j];
      --j;
    }
    d.inc[j + 1] = v;
  }
}

// InteractionManager::ZeroDrTot -> Mesh/Object::ZeroDrTot
// (src/interaction_manager.cpp:316-321, src/mesh.cpp:546-551).
__global__ void k_zero_dr(Dev d, Geo g) {
  if (guard_off(d.ctl, kGuardRebuild)) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) d.ctl->n_rebuilds += 1;
  if (i >= g.ns) return;
  for (int q = 0; q < 3; ++q) d.dr0[q * g.ns + i] = d.r[q * g.ns + i];
  d.acc[i] = 0.0;
}

static inline unsigned nblocks(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  return (unsigned)(b > 0 ? b : 1);
}

void launch_rebuild(const Dev &d, const Geo &g, hipStream_t st) {
  const int bs = 256;
  unsigned gb = nblocks(g.nb, bs);
  hipLaunchKernelGGL(k_keys, dim3(gb), dim3(bs), 0, st, d, g);
  launch_scan(d.cell_count, d.cell_start, g.n_cells, d.scan_tmp, d.ctl,
              kGuardRebuild, st);
  hipLaunchKernelGGL(k_scatter, dim3(gb), dim3(bs), 0, st, d, g);
  hipLaunchKernelGGL(k_run_sort, dim3(gb), dim3(bs), 0, st, d, g);
  hipLaunchKernelGGL(k_count, dim3(gb), dim3(bs), 0, st, d, g);
  launch_scan(d.ccount, d.coff, g.nb, d.scan_tmp, d.ctl, kGuardRebuild, st);
  hipLaunchKernelGGL(k_check_cap, dim3(1), dim3(64), 0, st, d, g);
  hipLaunchKernelGGL(k_emit, dim3(gb), dim3(bs), 0, st, d, g,
                     d.slot64 ? (int)SC_PREC_FP64 : (int)SC_PREC_MIXED);
  launch_scan(d.deg, d.inc_start, g.nb, d.scan_tmp, d.ctl, kGuardRebuild, st);
  hipLaunchKernelGGL(k_inc_fill, dim3(2048), dim3(bs), 0, st, d, g);
  hipLaunchKernelGGL(k_inc_sort, dim3(gb), dim3(bs), 0, st, d, g);
  hipLaunchKernelGGL(k_zero_dr, dim3(nblocks(g.ns, bs)), dim3(bs), 0, st, d, g);
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
    atomicMax((unsigned long long *)&d.ctl->dr_bits,
              (unsigned long long)__double_as_longlong(a));
}

void launch_redo_dr(const Dev &d, const Geo &g, hipStream_t st) {
  hipLaunchKernelGGL(k_redo_dr, dim3(nblocks(g.ns, 256)), dim3(256), 0, st, d, g);
}

void launch_decide(const Dev &d, const Geo &g, hipStream_t st) {
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, st, d.ctl, g.thr);
}

void launch_commit(const Dev &d, hipStream_t st) {
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(64), 0, st, d.ctl);
}

void launch_zero_i32(int32_t *p, int64_t n, hipStream_t st) {
  unsigned gb = nblocks(n, 256);
  if (gb > 4096) gb = 4096;
  hipLaunchKernelGGL(k_zero_i32, dim3(gb), dim3(256), 0, st, p, n);
}

// ---------------------------------------------------------------------------
// Pair pass
// ---------------------------------------------------------------------------

// ProcessPairInteraction for every live candidate
// (src/interaction_manager.cpp:323-397): minimum distance, then the
// potential only when dr^2 <= rc^2; the slot keeps its previous force
// otherwise.  lambda/mu (the contacts along u1/u2) are stored for the
// torque cross products done in the reduction.
__global__ __launch_bounds__(256) void k_pair64(Dev d, Geo g) {
  if (d.ctl->abort) return;
  const int64_t nc = d.ctl->nc;
  const int64_t nb = g.nb;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += st) {
    int2 pr = d.pairs[c];
    const int32_t b1 = pr.x, b2 = pr.y;
    double s1[3] = {d.spos[b1], d.spos[nb + b1], d.spos[2 * nb + b1]};
    double s2[3] = {d.spos[b2], d.spos[nb + b2], d.spos[2 * nb + b2]};
    double u1[3] = {d.u[b1], d.u[nb + b1], d.u[2 * nb + b1]};
    double u2[3] = {d.u[b2], d.u[nb + b2], d.u[2 * nb + b2]};
    double dr[3], dr2, lam, mu;
    sphero64(g.h, s1, u1, d.len[b1], s2, u2, d.len[b2], dr, &dr2, &lam, &mu);
    if (!(dr2 > g.rcut2)) {
      double f[3];
      if (pair_force64(g.c12, g.c6, g.fcut, dr, dr2, lam, mu, u1, u2, f))
        d.slot64[c] = make_double4(f[0], f[1], f[2], 0.0);
    }
    d.lm64[c] = make_double2(lam, mu);
  }
}

// ApplyPairInteractions (src/interaction_manager.cpp:500-517) as a
// deterministic per-bond segmented sum over the incidence CSR, in candidate
// order: obj1 gets +F and +contact1 x F, obj2 gets -F and -contact2 x F
// (torques as at src/interaction_manager.cpp:469-470).
__global__ __launch_bounds__(256) void k_reduce64(Dev d, Geo g) {
  if (d.ctl->abort) return;
  const int64_t nb = g.nb;
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double ub[3] = {d.u[b], d.u[nb + b], d.u[2 * nb + b]};
  double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
  const int32_t s = d.inc_start[b], e = d.inc_start[b + 1];
  for (int32_t q = s; q < e; ++q) {
    const int32_t ent = d.inc[q];
    const int32_t c = ent >> 1;
    const double4 fs = d.slot64[c];
    const double fc[3] = {fs.x, fs.y, fs.z};
    const double2 lm = d.lm64[c];
    double con[3], t[3];
    if ((ent & 1) == 0) {
      for (int k = 0; k < 3; ++k) con[k] = lm.x * ub[k];
      cross3(con, fc, t);
      for (int k = 0; k < 3; ++k) F[k] += fc[k];
      for (int k = 0; k < 3; ++k) T[k] += t[k];
    } else {
      for (int k = 0; k < 3; ++k) con[k] = lm.y * ub[k];
      cross3(con, fc, t);
      for (int k = 0; k < 3; ++k) F[k] -= fc[k];
      for (int k = 0; k < 3; ++k) T[k] -= t[k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    d.bf[k * nb + b] = F[k];
    d.bt[k * nb + b] = T[k];
  }
}

// ---------------------------------------------------------------------------
// Bond state from site positions: Bond::ReInit + Object::UpdatePeriodic
// (src/bond.cpp:22-37, src/object.cpp:202-210,
// src/linear_algebra.cpp:100-117).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bond_reinit_dev(const Dev &d, const Geo &g,
                                                int64_t b, const double *r1,
                                                const double *r2, int prec) {
  double o[3], l = 0.0;
  for (int k = 0; k < 3; ++k) {
    o[k] = r2[k] - r1[k];
    l += o[k] * o[k];
  }
  l = sqrt(l);
  double pos[3];
  for (int k = 0; k < 3; ++k) {
    pos[k] = r1[k] + 0.5 * o[k];
    o[k] /= l;
  }
  const int64_t nb = g.nb;
  for (int k = 0; k < 3; ++k) {
    double si = 0.0;
    for (int j = 0; j < 3; ++j) si += g.hinv[3 * k + j] * pos[j];
    si -= (double)nint(si);
    d.spos[k * nb + b] = si;
    d.u[k * nb + b] = o[k];
    d.pos[k * nb + b] = pos[k];
  }
  d.len[b] = l;
  if (prec == SC_PREC_MIXED)
    d.ub32[b] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)(0.5 * l));
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
    bond_reinit_dev(d, g, b0 + i, r1, r2, prec);
  }
}

void launch_bond_init(const Dev &d, const Geo &g, int precision,
                      hipStream_t st) {
  dim3 bs(64, 4);
  hipLaunchKernelGGL(k_bond_init, dim3(nblocks(g.nf, 4)), bs, 0, st, d, g,
                     precision);
}

// ---------------------------------------------------------------------------
// Integrator: Filament::UpdatePosition (src/filament.cpp:394-428) for one
// filament per 64-lane wavefront.  Sites map to lanes; the tridiagonal
// solves, the metric determinant recursions and renormalisation are
// sequential in the reference and run on lane 0 over LDS.
// ---------------------------------------------------------------------------
template <int MAXN>
__global__ __launch_bounds__(64) void k_integrate(Dev d, Geo g,
                                                  const int32_t *list,
                                                  int prec) {
  const Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  const int f = list[blockIdx.x];
  const int lane = threadIdx.x;
  const int64_t s0 = d.site_off[f];
  const int n = (int)(d.site_off[f + 1] - s0);
  const int64_t b0 = s0 - f;
  const int64_t ns = g.ns, nb = g.nb;
  const bool mid = ctl->midstep != 0;
  const int64_t step = ctl->cur_step;
  const double *fp = d.fpar + 8 * f;
  const double bl = fp[0], kappa = fp[1], gpar = fp[2], gperp = fp[3];
  const double sgpar = fp[4], sgperp = fp[5];

  __shared__ double u[3][MAXN], fs[3][MAXN], tn[3][MAXN], fr[3][MAXN];
  __shared__ double rn[3][MAXN], uo[3][MAXN];
  __shared__ double cs[MAXN], ta[MAXN], tb[MAXN], tc[MAXN], td[MAXN], ke[MAXN];
  __shared__ double dt[MAXN + 1], db[MAXN + 1];

  for (int i = lane; i < n - 1; i += 64)
    for (int k = 0; k < 3; ++k) u[k][i] = d.u[k * nb + b0 + i];
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < 3; ++k) u[k][n - 1] = u[k][n - 2];
  __syncthreads();

  // ApplyInteractionForces (src/filament.cpp:992-1048), driving off.
  const double linv = 1.0 / bl;
  for (int i = lane; i < n; i += 64) {
    double fsum[3] = {0.0, 0.0, 0.0};
    if (i >= 1) {
      const int64_t j = b0 + i - 1;
      double F[3] = {d.bf[j], d.bf[nb + j], d.bf[2 * nb + j]};
      double T[3] = {d.bt[j], d.bt[nb + j], d.bt[2 * nb + j]};
      double uj[3] = {u[0][i - 1], u[1][i - 1], u[2][i - 1]};
      double pt[3], sf[3];
      cross3(uj, T, pt);
      for (int k = 0; k < 3; ++k) {
        pt[k] *= linv;
        sf[k] = 0.5 * F[k];
      }
      for (int k = 0; k < 3; ++k) fsum[k] += sf[k];
      for (int k = 0; k < 3; ++k) pt[k] *= -1;
      for (int k = 0; k < 3; ++k) fsum[k] += pt[k];
    }
    if (i <= n - 2) {
      const int64_t j = b0 + i;
      double F[3] = {d.bf[j], d.bf[nb + j], d.bf[2 * nb + j]};
      double T[3] = {d.bt[j], d.bt[nb + j], d.bt[2 * nb + j]};
      double uj[3] = {u[0][i], u[1][i], u[2][i]};
      double pt[3], sf[3];
      cross3(uj, T, pt);
      for (int k = 0; k < 3; ++k) {
        pt[k] *= linv;
        sf[k] = 0.5 * F[k];
      }
      for (int k = 0; k < 3; ++k) fsum[k] += sf[k];
      for (int k = 0; k < 3; ++k) fsum[k] += pt[k];
    }
    for (int k = 0; k < 3; ++k) fs[k][i] = fsum[k];
  }
  // CalculateAngles / CalculateTangents (src/filament.cpp:430-443,
  // src/site.cpp:88-102)
  for (int i = lane; i < n; i += 64) {
    if (i < n - 2) {
      double a[3] = {u[0][i], u[1][i], u[2][i]};
      double b[3] = {u[0][i + 1], u[1][i + 1], u[2][i + 1]};
      cs[i] = dot3(a, b);
    }
    if (i == 0 || i == n - 1) {
      int bb = (i == 0 ? 0 : n - 2);
      for (int k = 0; k < 3; ++k) tn[k][i] = u[k][bb];
    } else {
      double t[3];
      for (int k = 0; k < 3; ++k) t[k] = u[k][i - 1] + u[k][i];
      normalize3(t);
      for (int k = 0; k < 3; ++k) tn[k][i] = t[k];
    }
  }
  __syncthreads();

  if (mid) {
    if (g.stoch) {
      // ConstructUnprojectedRandomForces (src/filament.cpp:518-548)
      const int64_t m = (step - 1) >> 1;
      for (int i = lane; i < n; i += 64) {
        double xi[3];
        noise_xi(g.seed, s0 + i, m, xi);
        const double t0 = tn[0][i], t1 = tn[1][i], t2 = tn[2][i];
        double xt[3];
        xt[0] = t0 * t0 * xi[0] + t0 * t1 * xi[1] + t0 * t2 * xi[2];
        xt[1] = t1 * t1 * xi[1] + t0 * t1 * xi[0] + t1 * t2 * xi[2];
        xt[2] = t2 * t2 * xi[2] + t0 * t2 * xi[0] + t1 * t2 * xi[1];
        for (int k = 0; k < 3; ++k) fr[k][i] = sgperp * xi[k] + (sgpar - sgperp) * xt[k];
      }
      __syncthreads();
      // GeometricallyProjectRandomForces (src/filament.cpp:550-599)
      for (int i = lane; i < n - 1; i += 64) {
        double tmp[3], ui[3] = {u[0][i], u[1][i], u[2][i]};
        for (int k = 0; k < 3; ++k) tmp[k] = fr[k][i + 1] - fr[k][i];
        td[i] = dot3(tmp, ui);
        tb[i] = 2;
        if (i > 0) {
          tc[i - 1] = -cs[i - 1];
          ta[i - 1] = -cs[i - 1];
        }
      }
      __syncthreads();
      if (lane == 0) thomas(n - 1, ta, tb, tc, td);
      __syncthreads();
      for (int i = lane; i < n; i += 64) {
        double p[3];
        if (i == 0) {
          for (int k = 0; k < 3; ++k) p[k] = fr[k][0] + td[0] * u[k][0];
        } else if (i == n - 1) {
          for (int k = 0; k < 3; ++k) p[k] = fr[k][i] - td[n - 2] * u[k][n - 2];
        } else {
          for (int k = 0; k < 3; ++k) p[k] = fr[k][i] + td[i] * u[k][i] - td[i - 1] * u[k][i - 1];
        }
        for (int k = 0; k < 3; ++k) d.fr[k * ns + s0 + i] = p[k];
      }
      __syncthreads();
      for (int i = lane; i < n; i += 64)
        for (int k = 0; k < 3; ++k) fr[k][i] = d.fr[k * ns + s0 + i];
    }
    // Mesh::UpdatePrevPositions (src/mesh.cpp:175-182)
    for (int i = lane; i < n; i += 64)
      for (int k = 0; k < 3; ++k) d.rp[k * ns + s0 + i] = d.r[k * ns + s0 + i];
  } else if (g.stoch) {
    for (int i = lane; i < n; i += 64)
      for (int k = 0; k < 3; ++k) fr[k][i] = d.fr[k * ns + s0 + i];
  }
  // AddRandomForces (src/filament.cpp:601-606)
  if (g.stoch)
    for (int i = lane; i < n; i += 64)
      for (int k = 0; k < 3; ++k) fs[k][i] += fr[k][i];

  // CalculateBendingForces, 3D (src/filament.cpp:608-776)
  if (lane == 0) {
    dt[0] = 1;
    dt[1] = 2;
    db[n] = 1;
    db[n - 1] = 2;
    for (int i = 2; i < n; ++i) {
      dt[i] = 2 * dt[i - 1] - (-cs[i - 2]) * (-cs[i - 2]) * dt[i - 2];
      db[n - i] = 2 * db[n - i + 1] - (-cs[n - i - 1]) * (-cs[n - i - 1]) * db[n - i + 2];
    }
  }
  __syncthreads();
  for (int i = lane; i < n - 2; i += 64) {
    double ginv = cs[i] * dt[i] * db[i + 3] / dt[n - 1];
    ke[i] = (kappa + bl * ginv) / (bl * bl);
  }
  __syncthreads();
  for (int kk = lane; kk < n; kk += 64) {
    double f3[3] = {0.0, 0.0, 0.0};
    if (kk > 1) {
      double u1[3] = {u[0][kk - 2], u[1][kk - 2], u[2][kk - 2]};
      double u2[3] = {u[0][kk - 1], u[1][kk - 1], u[2][kk - 1]};
      bend_add(f3, ke[kk - 2], u1, u2);
    }
    if (kk > 0 && kk < n - 1) {
      double u1[3] = {u[0][kk - 1], u[1][kk - 1], u[2][kk - 1]};
      double u2[3] = {u[0][kk], u[1][kk], u[2][kk]};
      bend_mid(f3, ke[kk - 1], u1, u2);
    }
    if (kk < n - 2) {
      double u1[3] = {u[0][kk], u[1][kk], u[2][kk]};
      double u2[3] = {u[0][kk + 1], u[1][kk + 1], u[2][kk + 1]};
      bend_sub(f3, ke[kk], u1, u2);
    }
    for (int k = 0; k < 3; ++k) fs[k][kk] += f3[k];
  }
  __syncthreads();

  // CalculateTensions (src/filament.cpp:778-838)
  const double igpar = 1.0 / gpar, igperp = 1.0 / gperp;
  for (int i = lane; i < n - 1; i += 64) {
    double t1[3] = {tn[0][i], tn[1][i], tn[2][i]};
    double t2[3] = {tn[0][i + 1], tn[1][i + 1], tn[2][i + 1]};
    double f1[3] = {fs[0][i], fs[1][i], fs[2][i]};
    double f2[3] = {fs[0][i + 1], fs[1][i + 1], fs[2][i + 1]};
    double fd[3];
    for (int a = 0; a < 3; ++a) {
      double ta_ = gam(igpar, igperp, t1, a, 0) * f1[0] + gam(igpar, igperp, t1, a, 1) * f1[1];
      ta_ += gam(igpar, igperp, t1, a, 2) * f1[2];
      double tb_ = gam(igpar, igperp, t2, a, 0) * f2[0] + gam(igpar, igperp, t2, a, 1) * f2[1];
      tb_ += gam(igpar, igperp, t2, a, 2) * f2[2];
      fd[a] = tb_ - ta_;
    }
    double u2[3] = {u[0][i], u[1][i], u[2][i]};
    td[i] = dot3(u2, fd);
    double a1 = dot3(t1, u2), a2 = dot3(t2, u2);
    tb[i] = 2.0 / gperp + (1.0 / gpar - 1.0 / gperp) * (a1 * a1 + a2 * a2);
    if (i > 0) {
      double u1[3] = {u[0][i - 1], u[1][i - 1], u[2][i - 1]};
      double hu = -1.0 / gperp * dot3(u2, u1) -
                  (1.0 / gpar - 1.0 / gperp) * (dot3(t1, u1) * dot3(t1, u2));
      tc[i - 1] = hu;
      ta[i - 1] = hu;
    }
  }
  __syncthreads();
  if (lane == 0) thomas(n - 1, ta, tb, tc, td);
  __syncthreads();
  for (int i = lane; i < n - 1; i += 64) d.tension[b0 + i] = td[i];

  // UpdateSitePositions (src/filament.cpp:840-927)
  const double delta = mid ? 0.5 * g.delta : g.delta;
  for (int i = lane; i < n; i += 64) {
    double f3[3] = {fs[0][i], fs[1][i], fs[2][i]};
    if (i == 0) {
      for (int k = 0; k < 3; ++k) f3[k] += td[0] * u[k][0];
    } else if (i == n - 1) {
      for (int k = 0; k < 3; ++k) f3[k] += -td[n - 2] * u[k][n - 2];
    } else {
      for (int k = 0; k < 3; ++k) f3[k] += td[i] * u[k][i] - td[i - 1] * u[k][i - 1];
    }
    double t3[3] = {tn[0][i], tn[1][i], tn[2][i]};
    for (int a = 0; a < 3; ++a) {
      double ft = gam(igpar, igperp, t3, a, 0) * f3[0] + gam(igpar, igperp, t3, a, 1) * f3[1];
      ft += gam(igpar, igperp, t3, a, 2) * f3[2];
      rn[a][i] = d.rp[a * ns + s0 + i] + delta * ft;
    }
  }
  __syncthreads();
  int bad = 0;
  for (int i = lane; i < n - 1; i += 64) {
    double rd[3], m = 0.0;
    for (int k = 0; k < 3; ++k) {
      rd[k] = rn[k][i + 1] - rn[k][i];
      m += rd[k] * rd[k];
    }
    m = sqrt(m);
    for (int k = 0; k < 3; ++k) uo[k][i] = rd[k] / m;
    // CheckBondLengths (src/filament.cpp:934-951)
    double diff = bl - m;
    double err = (diff < 0 ? -diff : diff) / bl;
    if (err > 1e-3) bad = 1;
  }
  const int renorm = __syncthreads_or(bad);
  if (lane == 0) {
    int32_t sw = d.norm[2 * f], nn = d.norm[2 * f + 1];
    if (renorm) {
      sw = !sw;
      nn = 0;
      if (sw) {
        for (int i = 1; i < n; ++i)
          for (int k = 0; k < 3; ++k) rn[k][i] = rn[k][i - 1] + bl * uo[k][i - 1];
      } else {
        for (int i = n - 1; i > 0; --i)
          for (int k = 0; k < 3; ++k) rn[k][i - 1] = rn[k][i] - bl * uo[k][i - 1];
      }
    }
    nn++;
    d.norm[2 * f] = sw;
    d.norm[2 * f + 1] = nn;
  }
  __syncthreads();

  // write back sites, then Mesh::UpdateBondPositions (src/mesh.cpp:324-334)
  for (int i = lane; i < n; i += 64)
    for (int k = 0; k < 3; ++k) d.r[k * ns + s0 + i] = rn[k][i];
  for (int i = lane; i < n - 1; i += 64) {
    double r1[3] = {rn[0][i], rn[1][i], rn[2][i]};
    double r2[3] = {rn[0][i + 1], rn[1][i + 1], rn[2][i + 1]};
    bond_reinit_dev(d, g, b0 + i, r1, r2, prec);
  }

  // Mesh::UpdateDrTot for the next step's rebuild test, after full steps
  // only, two Object::UpdateDrTot calls per site (src/mesh.cpp:533-544,
  // src/object.cpp:274-287).
  if (!mid) {
    double lmax = 0.0;
    for (int i = lane; i < n; i += 64) {
      double a = d.acc[s0 + i];
      d.acc_prev[s0 + i] = a;
      for (int rep = 0; rep < 2; ++rep)
        for (int k = 0; k < 3; ++k) {
          double dd = rn[k][i] - d.dr0[k * ns + s0 + i];
          a += dd * dd;
        }
      d.acc[s0 + i] = a;
      lmax = a > lmax ? a : lmax;
    }
    for (int o = 32; o > 0; o >>= 1) {
      double y = __shfl_xor(lmax, o, 64);
      lmax = y > lmax ? y : lmax;
    }
    if (lane == 0)
      atomicMax((unsigned long long *)&d.ctl->dr_bits,
                (unsigned long long)__double_as_longlong(lmax));
  }
}

void launch_integrate(const Dev &d, const Geo &g, const int32_t *bin_list,
                      int bin_count, int bin_maxn, int precision,
                      hipStream_t st) {
  if (bin_count <= 0) return;
  dim3 grid(bin_count), block(64);
  switch (bin_maxn) {
    case 32:
      hipLaunchKernelGGL(k_integrate<32>, grid, block, 0, st, d, g, bin_list, precision);
      break;
    case 64:
      hipLaunchKernelGGL(k_integrate<64>, grid, block, 0, st, d, g, bin_list, precision);
      break;
    case 128:
      hipLaunchKernelGGL(k_integrate<128>, grid, block, 0, st, d, g, bin_list, precision);
      break;
    case 256:
      hipLaunchKernelGGL(k_integrate<256>, grid, block, 0, st, d, g, bin_list, precision);
      break;
    default:
      hipLaunchKernelGGL(k_integrate<512>, grid, block, 0, st, d, g, bin_list, precision);
      break;
  }
}

void launch_pair(const Dev &d, const Geo &g, int precision, int grid,
                 hipStream_t st) {
  if (precision == SC_PREC_FP64) {
    hipLaunchKernelGGL(k_pair64, dim3(grid), dim3(256), 0, st, d, g);
  } else {
    launch_pair32(d, g, grid, st);
  }
}

void launch_reduce(const Dev &d, const Geo &g, int precision, hipStream_t st) {
  if (precision == SC_PREC_FP64) {
    hipLaunchKernelGGL(k_reduce64, dim3(nblocks(g.nb, 256)), dim3(256), 0, st, d, g);
  } else {
    launch_reduce32(d, g, st);
  }
}

}  // namespace sc
