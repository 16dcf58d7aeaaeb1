// sc_crosslink.hip -- harmonic crosslinks between bonds.
//
// Force law and application: Crosslink::CalculateTetherForces
// (src/crosslink.cpp:246-273) with MinimumDistance::PointPoint
// (src/minimum_distance.cpp:71-101) between anchor points placed by
// Anchor::UpdateAnchorPositionToBond (src/anchor.cpp:236-248); torques of
// Anchor::ApplyAnchorForces (src/anchor.cpp:132-146).  Even steps recompute
// the forces (k_xl_force); every step the pair kernel adds the stored
// forces to the bonds first, in crosslink order, from per-bond lists
// {crosslink << 1 | anchor} sorted by crosslink (the reference's serial
// application order, src/crosslink_species.cpp:262-266).  Rebinding events
// (k_xl_events, even steps) and the list rebuild they trigger run after the
// pair kernel; the lists are rebuilt only when an event happened
// (Ctl::xl_dirty), on the device, so the step stays graph-replayable.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

namespace {

// anchor point, orientation and length of bond b from its sites, with the
// arithmetic of Bond::ReInit (src/bond.cpp:22-37)
__device__ __forceinline__ void anchor_point(const Dev &d, const Geo &g, int32_t b, double lam,
                                             double *P, double *u, double *len) {
  const int64_t site = (int64_t)b + d.bond_fil[b];
  const int64_t ns = g.ns;
  double o[3], l = 0.0, pos[3];
  for (int k = 0; k < 3; ++k) {
    o[k] = d.r[k * ns + site + 1] - d.r[k * ns + site];
    l += o[k] * o[k];
  }
  l = sqrt(l);
  for (int k = 0; k < 3; ++k) {
    pos[k] = d.r[k * ns + site] + 0.5 * o[k];
    o[k] /= l;
  }
  for (int k = 0; k < 3; ++k) {
    u[k] = o[k];
    P[k] = pos[k] - (0.5 * l - lam) * u[k];
  }
  *len = l;
}

__device__ __forceinline__ void scaled(const Geo &g, const double *P, double *sp) {
  for (int i = 0; i < 3; ++i) {
    double si = 0.0;
    for (int j = 0; j < 3; ++j) si += g.hinv[3 * i + j] * P[j];
    si -= (double)nint(si);
    sp[i] = si;
  }
}

// scaled centre, orientation and length of bond b from its sites with the
// arithmetic of the FP64 bond record (Bond::ReInit + Object::UpdatePeriodic,
// src/bond.cpp:22-37, src/object.cpp:202-210): the partner search reads
// these rather than the record, which on the mixed path is compact (FP32)
__device__ __forceinline__ void bond_state(const Dev &d, const Geo &g, int32_t b, double *sp,
                                           double *u, double *len) {
  const int64_t site = (int64_t)b + d.bond_fil[b];
  const int64_t ns = g.ns;
  double o[3], l = 0.0, pos[3];
  for (int k = 0; k < 3; ++k) {
    o[k] = d.r[k * ns + site + 1] - d.r[k * ns + site];
    l += o[k] * o[k];
  }
  l = sqrt(l);
  for (int k = 0; k < 3; ++k) {
    pos[k] = d.r[k * ns + site] + 0.5 * o[k];
    u[k] = o[k] / l;
  }
  *len = l;
  scaled(g, pos, sp);
}

__device__ __forceinline__ double unit_draw(uint32_t v) {
  return ((double)v + 0.5) * 2.3283064365386962890625e-10;
}

__global__ __launch_bounds__(256) void k_xl_force(Dev d, Geo g) {
  if (d.ctl->abort) return;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.nx) return;
  double P[2][3], u[2][3], len[2], sp[2][3];
  for (int a = 0; a < 2; ++a) {
    anchor_point(d, g, d.xb[2 * x + a], d.xlam[2 * x + a], P[a], u[a], &len[a]);
    scaled(g, P[a], sp[a]);
  }
  double ds[3], dr[3], dr2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    ds[i] = sp[1][i] - sp[0][i];
    ds[i] -= (double)nint(ds[i]);
  }
  for (int i = 0; i < 3; ++i) {
    dr[i] = 0.0;
    for (int j = 0; j < 3; ++j) dr[i] += g.h[3 * i + j] * ds[j];
  }
  for (int i = 0; i < 3; ++i) dr2 += dr[i] * dr[i];
  const double length = sqrt(dr2);
  const double stretch = length - g.xr0;
  double tf;
  if (g.xkc >= 0 && stretch < 0)
    tf = g.xkc * stretch;
  else
    tf = g.xk * stretch;
  double F[3];
  for (int i = 0; i < 3; ++i) F[i] = tf * (dr[i] / length);
  double *o = d.xf + 9 * x;
  for (int i = 0; i < 3; ++i) o[i] = F[i];
  for (int a = 0; a < 2; ++a) {
    double dl[3], Fa[3], T[3];
    for (int i = 0; i < 3; ++i) {
      dl[i] = (d.xlam[2 * x + a] - 0.5 * len[a]) * u[a][i];
      Fa[i] = a == 0 ? F[i] : -F[i];
    }
    cross3(dl, Fa, T);
    for (int i = 0; i < 3; ++i) o[3 + 3 * a + i] = T[i];
  }
}

__global__ __launch_bounds__(256) void k_xl_events(Dev d, Geo g) {
  Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.nx) return;
  const int64_t m = ctl->cur_step / 2;
  uint32_t o[4] = {(uint32_t)x, (uint32_t)((uint64_t)x >> 32), (uint32_t)m, 1u};
  const uint32_t k0 = (uint32_t)g.xseed, k1 = (uint32_t)(g.xseed >> 32);
  philox4x32(o, k0, k1);
  if (!(unit_draw(o[0]) < g.xpoff)) return;
  const int am = unit_draw(o[1]) < 0.5 ? 0 : 1, as = 1 - am;
  const int32_t bs = d.xb[2 * x + as];
  double P[3], u[3], len, sP[3];
  anchor_point(d, g, bs, d.xlam[2 * x + as], P, u, &len);
  scaled(g, P, sP);
  const int n = g.n1d, R = g.xreach;
  const int32_t key0 = cell_key(n, sP[0], sP[1], sP[2]);
  const int cx = key0 / (n * n), cy = (key0 / n) % n, cz = key0 % n;
  const int span = 2 * R + 1 >= n ? n : 2 * R + 1;
  const int off0 = 2 * R + 1 >= n ? 0 : -R;
  int32_t best = -1, best_caller = 0;
  uint32_t best_score = 0;
  for (int dx = 0; dx < span; ++dx)
    for (int dy = 0; dy < span; ++dy)
      for (int dz = 0; dz < span; ++dz) {
        const int X = ((cx + off0 + dx) % n + n) % n, Y = ((cy + off0 + dy) % n + n) % n;
        const int Z = ((cz + off0 + dz) % n + n) % n;
        if (!cell_occupied(d, g, X, Y, Z)) continue;  // empty cell
        const int2 run = d.cell_run[(X * n + Y) * n + Z];
        for (int32_t j = run.x; j < run.y; ++j) {
          const int32_t b = d.sorted[j];
          if (b == bs) continue;
          double rec[7];
          bond_state(d, g, b, rec, rec + 3, rec + 6);
          double ds[3], dr[3], t = 0.0, d2 = 0.0;
          for (int i = 0; i < 3; ++i) {
            ds[i] = rec[i] - sP[i];
            ds[i] -= (double)nint(ds[i]);
          }
          for (int i = 0; i < 3; ++i) {
            dr[i] = 0.0;
            for (int k = 0; k < 3; ++k) dr[i] += g.h[3 * i + k] * ds[k];
          }
          for (int i = 0; i < 3; ++i) t += dr[i] * rec[3 + i];
          const double hl = 0.5 * rec[6];
          if (t > hl) t = hl;
          if (t < -hl) t = -hl;
          for (int i = 0; i < 3; ++i) {
            const double e = dr[i] - t * rec[3 + i];
            d2 += e * e;
          }
          if (!(d2 <= g.xrb2)) continue;
          const int32_t bc = d.bond_caller[b];
          uint32_t q[4] = {(uint32_t)x, (uint32_t)m, (uint32_t)bc, 2u};
          philox4x32(q, k0, k1);
          if (best < 0 || q[0] < best_score || (q[0] == best_score && bc < best_caller)) {
            best = b;
            best_caller = bc;
            best_score = q[0];
          }
        }
      }
  if (best < 0) return;
  d.xb[2 * x + am] = best;
  d.xlam[2 * x + am] = unit_draw(o[2]) * d.fpar[8 * d.bond_fil[best]];
  for (int i = 0; i < 9; ++i) d.xf[9 * x + i] = 0.0;
  atomicAdd((unsigned long long *)&ctl->xl_events, 1ull);
  ctl->xl_dirty = 1;
}

__device__ __forceinline__ bool xl_off(const Ctl *c) { return c->abort || !c->xl_dirty; }

__global__ __launch_bounds__(256) void k_xl_count(Dev d, Geo g) {
  if (xl_off(d.ctl)) return;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.nx) return;
  atomicAdd(&d.xdeg[d.xb[2 * x]], 1);
  atomicAdd(&d.xdeg[d.xb[2 * x + 1]], 1);
}

// places the entries; xdeg returns to zero for the next rebuild
__global__ __launch_bounds__(256) void k_xl_fill(Dev d, Geo g) {
  if (xl_off(d.ctl)) return;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.nx) return;
  for (int a = 0; a < 2; ++a) {
    const int32_t b = d.xb[2 * x + a];
    const int32_t pos = d.xinc_start[b] + atomicSub(&d.xdeg[b], 1) - 1;
    d.xinc[pos] = (int32_t)((x << 1) | a);
  }
}

__global__ __launch_bounds__(256) void k_xl_sort(Dev d, Geo g) {
  if (xl_off(d.ctl)) return;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nb) return;
  const int32_t q0 = d.xinc_start[b], q1 = d.xinc_start[b + 1];
  for (int32_t i = q0 + 1; i < q1; ++i) {
    const int32_t v = d.xinc[i];
    int32_t j = i - 1;
    while (j >= q0 && d.xinc[j] > v) {
      d.xinc[j + 1] = d.xinc[j];
      --j;
    }
    d.xinc[j + 1] = v;
  }
}

__global__ void k_xl_clean(Ctl *c) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !c->abort) c->xl_dirty = 0;
}

inline dim3 grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return dim3((unsigned)(b > 0 ? b : 1));
}

}  // namespace

void launch_xlink_forces(const Dev &d, const Geo &g, hipStream_t st) {
  if (g.nx <= 0) return;
  hipLaunchKernelGGL(k_xl_force, grid_for(g.nx), dim3(256), 0, st, d, g);
}

void launch_xlink_events(const Dev &d, const Geo &g, hipStream_t st) {
  if (g.nx <= 0 || !(g.xpoff > 0)) return;
  hipLaunchKernelGGL(k_xl_events, grid_for(g.nx), dim3(256), 0, st, d, g);
}

void launch_xlink_lists(const Dev &d, const Geo &g, hipStream_t st) {
  if (g.nx <= 0) return;
  hipLaunchKernelGGL(k_xl_count, grid_for(g.nx), dim3(256), 0, st, d, g);
  launch_scan(d.xdeg, d.xinc_start, g.nb, d.scan_tmp, d.ctl, kGuardXlink, st);
  hipLaunchKernelGGL(k_xl_fill, grid_for(g.nx), dim3(256), 0, st, d, g);
  hipLaunchKernelGGL(k_xl_sort, grid_for(g.nb), dim3(256), 0, st, d, g);
  hipLaunchKernelGGL(k_xl_clean, dim3(1), dim3(64), 0, st, d.ctl);
}

}  // namespace sc
