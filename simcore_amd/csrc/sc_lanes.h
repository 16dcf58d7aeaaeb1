// sc_lanes.h -- lane-level building blocks of the register-resident
// integrators: neighbour exchange across lanes, parallel cyclic reduction,
// 2x2 matrix and sum scans over a group of lanes, and the reduced-form
// vector helpers of the fast integrator.
//
// Lane exchange: a shift by a compile-time distance D is made of VALU lane
// moves, with no LDS round trip: DPP wave_shr:1 / wave_shl:1 across the
// whole wavefront; DPP row_shr / row_shl inside 16-lane rows (enough for
// groups of at most 16 lanes, which sit inside one row); for groups of 32
// or 64 lanes, the row shift for lanes whose source is in their own row and
// a row rotation carried to the neighbouring row by gfx950's permlane16 /
// permlane32 swaps for the others.  A lane whose source lies outside its
// group receives an unspecified value, which every caller replaces by its
// boundary value.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sc_math.h"

namespace sc {
namespace lanes {

// Keeps a permlane swap where it is written: the swap reads the source
// lanes that are active at that point, and a swap moved into a
// lane-dependent branch (the compiler may turn a select into one; unlike
// the DPP and ds_bpermute intrinsics, the swaps are not kept out of
// branches by a convergent attribute) would read inactive lanes.  An empty
// volatile asm cannot be moved into a branch, so the swap before it runs
// with every lane of the caller active.
__device__ __forceinline__ int pin(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// DPP move of a 64-bit value (two 32-bit moves); lanes without a valid
// source read 0 (no old value to preserve, so the move writes a fresh
// register instead of a copy of its source)
template <int CTRL>
__device__ __forceinline__ double dpp64(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
  const int nlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
  const int nhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)nhi << 32) | (uint32_t)nlo));
}

// 32-bit row moves of gfx950's permlane swaps (16-lane rows r0..r3):
// down16: rows 1, 3 receive rows 0, 2; up16: rows 0, 2 receive rows 1, 3;
// down32: rows 2, 3 receive rows 0, 1; up32: rows 0, 1 receive rows 2, 3.
// The other rows receive unspecified values.
__device__ __forceinline__ int down16(int v) {
  return pin((int)__builtin_amdgcn_permlane16_swap(v, v, false, false)[0]);
}
__device__ __forceinline__ int up16(int v) {
  return pin((int)__builtin_amdgcn_permlane16_swap(v, v, false, false)[1]);
}
__device__ __forceinline__ int down32(int v) {
  return pin((int)__builtin_amdgcn_permlane32_swap(v, v, false, false)[0]);
}
__device__ __forceinline__ int up32(int v) {
  return pin((int)__builtin_amdgcn_permlane32_swap(v, v, false, false)[1]);
}

// row k receives row k-1 (DOWN) or row k+1 (!DOWN) for every row whose
// source row is in the same group of G = 32 or 64 lanes
template <int G, bool DOWN>
__device__ __forceinline__ int row_step(int v, int lane) {
  const int row = lane >> 4;
  if constexpr (DOWN) {
    const int b = down16(v);
    if constexpr (G == 32) return b;
    const int c = down32(up16(v));  // row 2 <- row 1 via row 0
    return row == 2 ? c : b;
  } else {
    const int b = up16(v);
    if constexpr (G == 32) return b;
    const int c = up32(down16(v));  // row 1 <- row 2 via row 3
    return row == 1 ? c : b;
  }
}

// 32-bit value of lane + D inside groups of G = 32 or 64 lanes, 1 < |D| < 64
template <int G, int D>
__device__ __forceinline__ int wide_shift32(int v, int lane) {
  constexpr int s = D < 0 ? -D : D;
  if constexpr (s == 32) {
    return D < 0 ? down32(v) : up32(v);
  } else if constexpr (s == 16) {
    return row_step<G, (D < 0)>(v, lane);
  } else if constexpr (D < 0) {
    // inside the row: row_shr:s; the first s lanes of a row: row_ror:s
    // then the row below
    const int a = __builtin_amdgcn_mov_dpp(v, 0x110 + s, 0xF, 0xF, true);
    const int r = __builtin_amdgcn_mov_dpp(v, 0x120 + s, 0xF, 0xF, true);
    const int b = row_step<G, true>(r, lane);
    return (lane & 15) >= s ? a : b;
  } else {
    // inside the row: row_shl:s; the last s lanes of a row: row_ror:16-s
    // then the row above
    const int a = __builtin_amdgcn_mov_dpp(v, 0x100 + s, 0xF, 0xF, true);
    const int r = __builtin_amdgcn_mov_dpp(v, 0x120 + 16 - s, 0xF, 0xF, true);
    const int b = row_step<G, false>(r, lane);
    return (lane & 15) < 16 - s ? a : b;
  }
}

// value of lane + D for lanes whose source is in their own group of G
// lanes (D compile-time)
template <int G, int D>
__device__ __forceinline__ double lane_shift(double v, int lane) {
  static_assert(D > -64 && D < 64, "shift within a wavefront");
  if constexpr (D == 0) {
    return v;
  } else if constexpr (D == -1) {
    return dpp64<0x138>(v);  // wave_shr:1
  } else if constexpr (D == 1) {
    return dpp64<0x130>(v);  // wave_shl:1
  } else if constexpr (G <= 16 && D < 0 && D > -16) {
    return dpp64<0x110 + (-D)>(v);  // row_shr:-D
  } else if constexpr (G <= 16 && D > 0 && D < 16) {
    return dpp64<0x100 + D>(v);  // row_shl:D
  } else if constexpr (G == 32 || (G == 64 && (D == 32 || D == -32))) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const int lo = wide_shift32<G, D>((int)(uint32_t)b, lane);
    const int hi = wide_shift32<G, D>((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
  } else {
    return __shfl(v, (lane + D) & 63, 64);
  }
}

// lane_shift with a distance that is a constant once loops are unrolled
template <int G>
__device__ __forceinline__ double lane_shift_c(double v, int lane, int d) {
  switch (d) {
    case 0: return v;
    case -1: return lane_shift<G, -1>(v, lane);
    case 1: return lane_shift<G, 1>(v, lane);
    case -2: return lane_shift<G, -2>(v, lane);
    case 2: return lane_shift<G, 2>(v, lane);
    case -3: return lane_shift<G, -3>(v, lane);
    case 3: return lane_shift<G, 3>(v, lane);
    case -4: return lane_shift<G, -4>(v, lane);
    case 4: return lane_shift<G, 4>(v, lane);
    case -8: return lane_shift<G, -8>(v, lane);
    case 8: return lane_shift<G, 8>(v, lane);
    case -16: return lane_shift<G, -16>(v, lane);
    case 16: return lane_shift<G, 16>(v, lane);
    case -32: return lane_shift<G, -32>(v, lane);
    case 32: return lane_shift<G, 32>(v, lane);
    default: return __shfl(v, (lane + d) & 63, 64);
  }
}

// value held by lane-1 / lane+1 (lane 0 / lane 63 read 0)
__device__ __forceinline__ double d_prev(double v) { return dpp64<0x138>(v); }
__device__ __forceinline__ double d_next(double v) { return dpp64<0x130>(v); }
__device__ __forceinline__ void v_prev(const double *a, double *o) {
  for (int q = 0; q < 3; ++q) o[q] = d_prev(a[q]);
}
__device__ __forceinline__ void v_next(const double *a, double *o) {
  for (int q = 0; q < 3; ++q) o[q] = d_next(a[q]);
}

// value of group lane gl+delta; `fb` outside [0, G)
template <int G>
__device__ __forceinline__ double g_from(double v, int lane, int gl, int delta, double fb) {
  const int src = gl + delta;
  const bool ok = src >= 0 && src < G;
  const double x = lane_shift_c<G>(v, lane, delta);
  return ok ? x : fb;
}

__device__ __forceinline__ double dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void proj_perp(const double *a, const double *b, double *o) {
  const double ab = dot(a, b);
  for (int q = 0; q < 3; ++q) o[q] = a[q] - b[q] * ab;
}

__device__ __forceinline__ void mob(double igperp, double dg, const double *t,
                                    const double *f, double *o) {
  const double tf = dot(t, f);
  for (int q = 0; q < 3; ++q) o[q] = igperp * f[q] + dg * t[q] * tf;
}

// Parallel cyclic reduction of rows i < m (row i in lane gl = i): lower a
// (x[i-1]), diagonal b, upper c (x[i+1]), right-hand side d.  Rows >= m are
// inert.  Returns x[i].  Reciprocals by rcp_fast (mixed-precision path).
template <int G>
__device__ __forceinline__ double pcr(double a, double b, double c, double d, int m, int lane,
                                      int gl) {
  if (gl >= m) {
    a = 0.0;
    b = 1.0;
    c = 0.0;
    d = 0.0;
  }
#pragma unroll
  for (int s = 1; s < G; s <<= 1) {
    const double rb = rcp_fast(b);
    const double am = g_from<G>(a, lane, gl, -s, 0.0), cm = g_from<G>(c, lane, gl, -s, 0.0);
    const double dm = g_from<G>(d, lane, gl, -s, 0.0), rm = g_from<G>(rb, lane, gl, -s, 0.0);
    const double ap = g_from<G>(a, lane, gl, s, 0.0), cp = g_from<G>(c, lane, gl, s, 0.0);
    const double dp = g_from<G>(d, lane, gl, s, 0.0), rp = g_from<G>(rb, lane, gl, s, 0.0);
    const bool lo = gl - s >= 0, hi = gl + s < m;
    const double al = lo ? -a * rm : 0.0;
    const double ga = hi ? -c * rp : 0.0;
    const double na = al * am;
    const double nc = ga * cp;
    b = b + al * cm + ga * ap;
    d = d + al * dm + ga * dp;
    a = na;
    c = nc;
  }
  return d * rcp_fast(b);
}

// inclusive prefix (dir = -1: over lower lanes, newest on the left) or
// suffix (dir = +1: over higher lanes) product of 2x2 matrices
template <int G>
__device__ __forceinline__ void mat_scan(double *P, int lane, int gl, int dir) {
#pragma unroll
  for (int s = 1; s < G; s <<= 1) {
    double Q[4];
    const double id[4] = {1.0, 0.0, 0.0, 1.0};
    for (int k = 0; k < 4; ++k) Q[k] = g_from<G>(P[k], lane, gl, dir * s, id[k]);
    double R[4];
    // both directions: P covers the lanes nearer to this one, Q the farther
    // ones; prefix: P_i = M_i ... (P Q), suffix: S_j = N_j ... (P Q)
    R[0] = P[0] * Q[0] + P[1] * Q[2];
    R[1] = P[0] * Q[1] + P[1] * Q[3];
    R[2] = P[2] * Q[0] + P[3] * Q[2];
    R[3] = P[2] * Q[1] + P[3] * Q[3];
    for (int k = 0; k < 4; ++k) P[k] = R[k];
  }
}

// inclusive prefix (dir -1) / suffix (dir +1) sum over the group
template <int G>
__device__ __forceinline__ double sum_scan(double v, int lane, int gl, int dir) {
#pragma unroll
  for (int s = 1; s < G; s <<= 1) v += g_from<G>(v, lane, gl, dir * s, 0.0);
  return v;
}

}  // namespace lanes
}  // namespace sc
