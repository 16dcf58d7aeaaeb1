// sc_lanes.h -- lane-level building blocks of the register-resident
// integrators: neighbour exchange by lane shuffles (ds_bpermute), parallel
// cyclic reduction, 2x2 matrix and sum scans over a group of lanes, and the
// reduced-form vector helpers of the fast integrator.
#pragma once

#include <hip/hip_runtime.h>

namespace sc {
namespace lanes {

// value held by lane-1 / lane+1 (lane 0 / lane 63 read their own)
__device__ __forceinline__ double d_prev(double v) {
  const int l = (int)__lane_id();
  return __shfl(v, l > 0 ? l - 1 : 0, 64);
}
__device__ __forceinline__ double d_next(double v) {
  const int l = (int)__lane_id();
  return __shfl(v, l < 63 ? l + 1 : 63, 64);
}
__device__ __forceinline__ void v_prev(const double *a, double *o) {
  for (int q = 0; q < 3; ++q) o[q] = d_prev(a[q]);
}
__device__ __forceinline__ void v_next(const double *a, double *o) {
  for (int q = 0; q < 3; ++q) o[q] = d_next(a[q]);
}

// value of group lane gl+delta (any delta); `fb` outside [0, G)
template <int G>
__device__ __forceinline__ double g_from(double v, int lane, int gl, int delta, double fb) {
  const int src = gl + delta;
  const bool ok = src >= 0 && src < G;
  const double x = __shfl(v, ok ? lane + delta : lane, 64);
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
// inert.  Returns x[i].
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
    const double rb = 1.0 / b;
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
  return d / b;
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
