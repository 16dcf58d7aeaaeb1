// sc_math.h -- device math of the filament BD path, in the reference's
// operation order (the FP64 path is compared value-for-value with a CPU
// restatement; these functions must stay free of reassociation).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/simcore_amd.h"

namespace sc {

// 1/x for the mixed-precision path: the v_rcp_f64 estimate refined by one
// Newton step, e = 1 - x r, r += r e (the error squared).  The IEEE
// division sequence adds a second step, operand scaling and a final
// rounding fix-up, about three times the instructions; x here is a length,
// pivot or determinant of order one (0 gives a NaN instead of an infinity,
// in lanes whose values are unused).
__device__ __forceinline__ double rcp_fast(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) for the mixed-precision path (x a squared length of order
// one): the v_rsq_f64 estimate refined by one Newton step,
// r += r (1 - x r^2) / 2, in place of an IEEE square root and a reciprocal.
__device__ __forceinline__ double rsq_fast(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * r, r, 1.0);
  return fma(0.5 * r, e, r);
}

// NINT (include/simcore/macros.hpp:6): round half away from zero to int.
__device__ __forceinline__ int nint(double x) {
  return x < 0.0 ? (int)(x - 0.5) : (int)(x + 0.5);
}

// CellList::FindCellCoords (src/cell_list.cpp:139-157) packed as
// (x*n + y)*n + z, the order of MakePairs' cell_[x][y][z] walk.  Scaled
// coordinates lie in [-0.5, 0.5]; the clamp to [0, n-1] only guards the
// half-ulp case s = -0.5 - 2^-54, where the reference would index cell -1.
// cell coordinate of one scaled coordinate s, and n1d (s + 1/2) itself
__device__ __forceinline__ int cell_coord(int n1d, double s, double *t) {
  const double x = s + 0.5;
  *t = n1d * x;
  int xc = (int)floor(*t);
  if (xc == n1d) xc -= 1;
  return xc < 0 ? 0 : (xc >= n1d ? n1d - 1 : xc);
}

__device__ __forceinline__ int32_t cell_key(int n1d, double sx, double sy,
                                            double sz) {
  double t;
  const int x = cell_coord(n1d, sx, &t), y = cell_coord(n1d, sy, &t),
            z = cell_coord(n1d, sz, &t);
  return ((int32_t)x * n1d + y) * n1d + z;
}

// Mixed-precision bond record, 32 bytes (4 doubles of storage per bond),
// two uint4: {offset of the scaled position inside its cell, x, y, z, in
// units of 2^-32 cell; length, FP32 high part} and {unit vector in
// octahedral coordinates, two 32-bit fixed-point values over [-1, 1]; cell
// coordinates x | y << 10 | z << 20; length, FP32 low part}.  Decoded in
// FP64: the scaled position (c + offset) / n1d - 1/2 within 2^-32 of a cell
// (3e-10 sigma for cells of 1.13 sigma), the unit vector within about 1e-9,
// the length to about 1e-14 relative; the cell coordinates, hence the cell
// key, are those cell_key computes from the FP64 scaled position.  At most
// 1024 cells per side (checked at creation).
__host__ __device__ __forceinline__ uint32_t fix32(double f, double scale) {
  double v = floor(f * scale + 0.5);
  v = v < 0.0 ? 0.0 : (v > 4294967295.0 ? 4294967295.0 : v);
  return (uint32_t)v;
}

// octahedral map of a unit vector to [-1, 1]^2 (the lower hemisphere folded
// over the diagonals) and back
__host__ __device__ __forceinline__ void oct_encode(const double *u, uint32_t *ox,
                                                    uint32_t *oy) {
  const double a = fabs(u[0]) + fabs(u[1]) + fabs(u[2]);  // in [1, sqrt(3)]
#ifdef __HIP_DEVICE_COMPILE__
  const double ia = rcp_fast(a);
  double px = u[0] * ia, py = u[1] * ia;
#else
  double px = u[0] / a, py = u[1] / a;
#endif
  if (u[2] < 0.0) {
    const double qx = (1.0 - fabs(py)) * (px >= 0.0 ? 1.0 : -1.0);
    const double qy = (1.0 - fabs(px)) * (py >= 0.0 ? 1.0 : -1.0);
    px = qx;
    py = qy;
  }
  *ox = fix32(px + 1.0, 2147483648.0);
  *oy = fix32(py + 1.0, 2147483648.0);
}

__host__ __device__ __forceinline__ void oct_decode(uint32_t ox, uint32_t oy, double *u) {
  const double px = (double)ox * (1.0 / 2147483648.0) - 1.0;
  const double py = (double)oy * (1.0 / 2147483648.0) - 1.0;
  const double z = 1.0 - fabs(px) - fabs(py);
  const double x = z < 0.0 ? (1.0 - fabs(py)) * (px >= 0.0 ? 1.0 : -1.0) : px;
  const double y = z < 0.0 ? (1.0 - fabs(px)) * (py >= 0.0 ? 1.0 : -1.0) : py;
  const double n2 = x * x + y * y + z * z;  // in [1/3, 1]
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rsq(n2);
  r = r * (1.5 - 0.5 * n2 * r * r);
#else
  const double r = 1.0 / sqrt(n2);
#endif
  u[0] = x * r;
  u[1] = y * r;
  u[2] = z * r;
}

template <class GeoT>
__device__ __forceinline__ void rec_write_compact(double *br, const GeoT &g, int64_t b,
                                                  const double *sp, const double *u,
                                                  double l) {
  double t[3];
  int c[3];
  for (int k = 0; k < 3; ++k) c[k] = cell_coord(g.n1d, sp[k], &t[k]);
  uint32_t ox, oy;
  oct_encode(u, &ox, &oy);
  const float lh = (float)l;
  const float ll = (float)(l - (double)lh);
  uint4 *rec = reinterpret_cast<uint4 *>(br + 4 * b);
  rec[0] = make_uint4(fix32(t[0] - c[0], 4294967296.0), fix32(t[1] - c[1], 4294967296.0),
                      fix32(t[2] - c[2], 4294967296.0), __float_as_uint(lh));
  rec[1] = make_uint4(ox, oy, (uint32_t)(c[0] | (c[1] << 10) | (c[2] << 20)),
                      __float_as_uint(ll));
}

// cell key of a compact record (the key cell_key gives its scaled position)
template <class GeoT>
__device__ __forceinline__ int32_t rec_key_compact(const double *br, const GeoT &g,
                                                   int64_t b) {
  const uint4 v = reinterpret_cast<const uint4 *>(br + 4 * b)[1];
  const int packed = (int)v.z;
  const int x = packed & 1023, y = (packed >> 10) & 1023, z = (packed >> 20) & 1023;
  return ((int32_t)x * g.n1d + y) * g.n1d + z;
}

// scaled position s, unit vector u and length l of bond b: compact record
__device__ __forceinline__ void rec_read_compact(const double *br, double inv_n1d, int64_t b,
                                                 double *s, double *u, double *l) {
  const uint4 *rec = reinterpret_cast<const uint4 *>(br + 4 * b);
  const uint4 a = rec[0], v = rec[1];
  const int packed = (int)v.z;
  constexpr double k32 = 1.0 / 4294967296.0;
  s[0] = ((double)(packed & 1023) + (double)a.x * k32) * inv_n1d - 0.5;
  s[1] = ((double)((packed >> 10) & 1023) + (double)a.y * k32) * inv_n1d - 0.5;
  s[2] = ((double)((packed >> 20) & 1023) + (double)a.z * k32) * inv_n1d - 0.5;
  oct_decode(v.x, v.y, u);
  *l = (double)__uint_as_float(a.w) + (double)__uint_as_float(v.w);
}

// the same from a 64-byte FP64 record
__device__ __forceinline__ void rec_read_full(const double *br, int64_t b, double *s, double *u,
                                              double *l) {
  const double4 *rec = reinterpret_cast<const double4 *>(br + 8 * b);
  const double4 a = rec[0], v = rec[1];
  s[0] = a.x;
  s[1] = a.y;
  s[2] = a.z;
  u[0] = a.w;
  u[1] = v.x;
  u[2] = v.y;
  *l = v.z;
}

// either layout, as the system keeps it
template <class GeoT>
__device__ __forceinline__ void rec_read(const double *br, const GeoT &g, int64_t b, double *s,
                                         double *u, double *l) {
  if (g.compact)
    rec_read_compact(br, g.inv_n1d, b, s, u, l);
  else
    rec_read_full(br, b, s, u, l);
}

// Non-redundant neighbour stencil of AssignCellNeighbors(false)
// (src/cell_list.cpp:206-262) for a fully periodic 3D grid: (x+1,y,z);
// (x-1..x+1, y+1, z); (x-1..x+1, y-1..y+1, z+1), wrapped.
__device__ __forceinline__ void half_shell(int n, int32_t key, int32_t *out) {
  const int x = key / (n * n);
  const int y = (key / n) % n;
  const int z = key % n;
  int cnt = 0;
  for (int zp = z; zp < z + 2; ++zp) {
    int nz = (zp == n) ? 0 : zp;
    int y_begin = (zp > z ? y - 1 : y);
    for (int yp = y_begin; yp < y + 2; ++yp) {
      int ny = yp < 0 ? n - 1 : (yp == n ? 0 : yp);
      int x_begin = (yp > y || zp > z ? x - 1 : x + 1);
      for (int xp = x_begin; xp < x + 2; ++xp) {
        int nx = xp < 0 ? n - 1 : (xp == n ? 0 : xp);
        out[cnt++] = ((int32_t)nx * n + ny) * n + nz;
      }
    }
  }
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
  double m = 0.0;
  m += a[0] * b[0];
  m += a[1] * b[1];
  m += a[2] * b[2];
  return m;
}

__device__ __forceinline__ void cross3(const double *a, const double *b,
                                       double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void normalize3(double *a) {
  double m = 0.0;
  m += a[0] * a[0];
  m += a[1] * a[1];
  m += a[2] * a[2];
  m = sqrt(m);
  a[0] /= m;
  a[1] /= m;
  a[2] /= m;
}

// tridiagonal_solver (src/linear_algebra.cpp:164-197): Thomas, in place on
// c and d, n unknowns.
__device__ __forceinline__ void thomas(int n, const double *a, const double *b,
                                       double *c, double *d) {
  n--;
  c[0] /= b[0];
  d[0] /= b[0];
  for (int i = 1; i < n; ++i) {
    double den = b[i] - a[i - 1] * c[i - 1];
    c[i] /= den;
    d[i] = (d[i] - a[i - 1] * d[i - 1]) / den;
  }
  d[n] = (d[n] - a[n - 1] * d[n - 1]) / (b[n] - a[n - 1] * c[n - 1]);
  for (int i = n; i-- > 0;) d[i] -= c[i] * d[i + 1];
}

__device__ __forceinline__ double dabs(double x) { return x < 0 ? -x : x; }
__device__ __forceinline__ double dsign(double a, double b) {
  return b >= 0.0 ? fabs(a) : -fabs(a);
}

// MinimumDistance::Sphero for a fully periodic box
// (src/minimum_distance.cpp:202-344).  Returns r_min, |r_min|^2 and the
// contact parameters lambda (along u1) and mu (along u2).
__device__ __forceinline__ void sphero64(const double *h, const double *s_1,
                                         const double *u_1, double length_1,
                                         const double *s_2, const double *u_2,
                                         double length_2, double *r_min,
                                         double *r_min_mag2, double *lam_out,
                                         double *mu_out) {
  double dr[3], ds[3];
  const double hl1 = 0.5 * length_1, hl2 = 0.5 * length_2;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ds[i] = s_2[i] - s_1[i];
    ds[i] -= (double)nint(ds[i]);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    dr[i] = 0.0;
    for (int j = 0; j < 3; ++j) dr[i] += h[3 * i + j] * ds[j];
  }
  double a = 0.0, b = 0.0, c = 0.0;  // dr.u1, dr.u2, u1.u2
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a += dr[i] * u_1[i];
    b += dr[i] * u_2[i];
    c += u_1[i] * u_2[i];
  }
  double lambda, mu;
  const double denom = 1.0 - c * c;
  if (denom < 1.0e-12) {
    lambda = a / 2.0;
    mu = -b / 2.0;
  } else {
    lambda = (a - c * b) / denom;
    mu = (-b + c * a) / denom;
  }
  double lambda_mag = dabs(lambda), mu_mag = dabs(mu);
  if (lambda_mag > hl1 && mu_mag > hl2) {
    double la = dsign(hl1, lambda);
    double ma = -b + la * c;
    if (dabs(ma) > hl2) ma = dsign(hl2, ma);
    double ra[3], rb[3], m2a = 0.0, m2b = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ra[i] = dr[i] - la * u_1[i] + ma * u_2[i];
      m2a += ra[i] * ra[i];
    }
    double mb = dsign(hl2, mu);
    double lb = a + mb * c;
    if (dabs(lb) > hl1) lb = dsign(hl1, lb);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      rb[i] = dr[i] - lb * u_1[i] + mb * u_2[i];
      m2b += rb[i] * rb[i];
    }
    if (m2a < m2b) {
      lambda = la;
      mu = ma;
      *r_min_mag2 = m2a;
      for (int i = 0; i < 3; ++i) r_min[i] = ra[i];
    } else {
      lambda = lb;
      mu = mb;
      *r_min_mag2 = m2b;
      for (int i = 0; i < 3; ++i) r_min[i] = rb[i];
    }
  } else {
    if (lambda_mag > hl1) {
      lambda = dsign(hl1, lambda);
      mu = -b + lambda * c;
      if (dabs(mu) > hl2) mu = dsign(hl2, mu);
    } else if (mu_mag > hl2) {
      mu = dsign(hl2, mu);
      lambda = a + mu * c;
      if (dabs(lambda) > hl1) lambda = dsign(hl1, lambda);
    }
    double m2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      r_min[i] = dr[i] - lambda * u_1[i] + mu * u_2[i];
      m2 += r_min[i] * r_min[i];
    }
    *r_min_mag2 = m2;
  }
  *lam_out = lambda;
  *mu_out = mu;
}

// PotentialManager::CalcPotential (include/simcore/potential_manager.hpp:
// 92-103) -> MaxForcePotential when dr^2 < 1e-12
// (include/simcore/max_force_potential.hpp:11-40), else WCAPotential
// (include/simcore/wca_potential.hpp:14-32).  Returns 1 if force written.
// *rs receives the pair's signed distance: +|dr| on the WCA branch (the
// force is antiparallel to dr, so stress_ij = -dr_i F_j = (r/|F|) F_i F_j),
// -|dr| on the max-force branch (force parallel to dr).
__device__ __forceinline__ int pair_force64(double c12, double c6,
                                            double fcut, double *dr,
                                            double dr2, double lam, double mu,
                                            const double *u1, const double *u2,
                                            double *force, double *rs) {
  if (dr2 < 1e-12) {
    double c1[3], c2[3];
    for (int i = 0; i < 3; ++i) {
      c1[i] = lam * u1[i];
      c2[i] = mu * u2[i];
    }
    if (c1[0] != 0.0 || c1[1] != 0.0 || c1[2] != 0.0)
      for (int i = 0; i < 3; ++i) dr[i] = c1[i] - c2[i];
    double rmag = 0.0;
    for (int i = 0; i < 3; ++i) rmag += dr[i] * dr[i];
    rmag = sqrt(rmag);
    if (rmag < 1e-12) return 0;
    double rinv = 1.0 / (rmag);
    for (int i = 0; i < 3; ++i) force[i] = fcut * dr[i] * rinv;
    *rs = -rmag;
    return 1;
  }
  double rmag = sqrt(dr2);
  *rs = rmag;
  double rinv = 1.0 / (rmag);
  double rinv2 = rinv * rinv;
  double r6 = rinv2 * rinv2 * rinv2;
  double ffac = -(12.0 * c12 * r6 - 6.0 * c6) * r6 * rinv;
  if (dabs(ffac) > fcut) ffac = (ffac >= 0.0 ? 1 : -1) * fcut;
  for (int i = 0; i < 3; ++i) force[i] = ffac * dr[i] * rinv;
  return 1;
}

// Metric-corrected bending terms of Filament::CalculateBendingForces, 3D
// branch (src/filament.cpp:725-775).  For component d, p and q are the
// other two components in increasing order.
__device__ __forceinline__ double bterm(const double *a, const double *b,
                                        int d) {
  const int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
  return (1 - b[d] * b[d]) * a[d] - b[d] * b[p] * a[p] - b[d] * b[q] * a[q];
}
__device__ __forceinline__ void bend_add(double *f, double ke, const double *u1,
                                         const double *u2) {
  for (int d = 0; d < 3; ++d) f[d] += ke * bterm(u1, u2, d);
}
__device__ __forceinline__ void bend_mid(double *f, double ke, const double *u1,
                                         const double *u2) {
  for (int d = 0; d < 3; ++d) f[d] += ke * (bterm(u2, u1, d) - bterm(u1, u2, d));
}
__device__ __forceinline__ void bend_sub(double *f, double ke, const double *u1,
                                         const double *u2) {
  for (int d = 0; d < 3; ++d) f[d] -= ke * bterm(u2, u1, d);
}

// Element (a,b) of the site mobility 1/gpar t t^T + 1/gperp (I - t t^T)
// (src/filament.cpp:782-794).
__device__ __forceinline__ double gam(double igpar, double igperp,
                                      const double *t, int a, int b) {
  return igpar * (t[a] * t[b]) + igperp * ((a == b ? 1 : 0) - t[a] * t[b]);
}

// Philox4x32-10 (Salmon et al., SC'11): the injected thermal-noise stream.
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0,
                                           uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

// xi_d = (x_d + 1/2) 2^-32 - 1/2 (exact), standing in for
// RandomUniform() - 0.5 of src/filament.cpp:529-530; counter =
// (global site, midstep index).
__device__ __forceinline__ void noise_xi(uint64_t seed, int64_t site,
                                         int64_t m, double *xi) {
  uint32_t c[4] = {(uint32_t)site, (uint32_t)((uint64_t)site >> 32),
                   (uint32_t)m, (uint32_t)((uint64_t)m >> 32)};
  philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  for (int k = 0; k < 3; ++k)
    xi[k] = ((double)c[k] + 0.5) * 2.3283064365386962890625e-10 - 0.5;
}

// Potential energy of a pair at signed distance rs (pair_force64): the WCA
// value r6 (c12 r6 - c6) + eps with the reference's operations
// (include/simcore/wca_potential.hpp:15-31), 0 on the max-force branch
// (include/simcore/max_force_potential.hpp:38) or for a pair never inside
// the cutoff (rs = 0).
__device__ __forceinline__ double pair_pote(double c12, double c6, double eps, double rs) {
  if (!(rs > 0.0)) return 0.0;
  double rinv = 1.0 / (rs);
  double rinv2 = rinv * rinv;
  double r6 = rinv2 * rinv2 * rinv2;
  return r6 * (c12 * r6 - c6) + eps;
}

// The three draws of `site` (caller numbering) at midstep m: from the
// caller's buffer when one is set (zero outside the midsteps it holds; the
// host never steps past them), else from the counter-based stream.
template <class GeoT>
__device__ __forceinline__ void draw_xi(const GeoT &g, int64_t site, int64_t m, double *xi) {
  if (g.xi_buf) {
    const int64_t j = m - g.xi_m0;
    const bool ok = j >= 0 && j < g.xi_nm;
    const double *p = g.xi_buf + (ok ? (j * g.ns + site) * 3 : 0);
    for (int k = 0; k < 3; ++k) xi[k] = ok ? p[k] : 0.0;
  } else {
    noise_xi(g.seed, site, m, xi);
  }
}

// Bond::ReInit + Object::UpdatePeriodic (src/bond.cpp:22-37,
// src/object.cpp:202-210, src/linear_algebra.cpp:100-117) from the two site
// positions; writes the bond record {scaled position, unit orientation,
// length}: 64 bytes of FP64 on the FP64 path, the 32-byte compact record on
// the mixed one.  (Templated on the pointer/constant structs.)
template <class DevT, class GeoT>
__device__ __forceinline__ void bond_reinit(const DevT &d, const GeoT &g,
                                            int64_t b, const double *r1,
                                            const double *r2, int prec) {
  double o[3], l = 0.0;
  for (int k = 0; k < 3; ++k) {
    o[k] = r2[k] - r1[k];
    l += o[k] * o[k];
  }
  l = sqrt(l);
  // mixed precision: u = o (1/l), the arithmetic of the register integrators,
  // which recompute u from the site positions
  const double il = prec == SC_PREC_MIXED ? rcp_fast(l) : 0.0;
  double pos[3];
  for (int k = 0; k < 3; ++k) {
    pos[k] = r1[k] + 0.5 * o[k];
    if (prec == SC_PREC_MIXED)
      o[k] *= il;
    else
      o[k] /= l;
  }
  double sp[3];
  for (int k = 0; k < 3; ++k) {
    double si = 0.0;
    for (int j = 0; j < 3; ++j) si += g.hinv[3 * k + j] * pos[j];
    si -= (double)nint(si);
    sp[k] = si;
  }
  if (prec == SC_PREC_MIXED) {
    rec_write_compact(d.br_out, g, b, sp, o, l);
    return;
  }
  double4 *rec = reinterpret_cast<double4 *>(d.br_out + 8 * b);
  rec[0] = make_double4(sp[0], sp[1], sp[2], o[0]);
  rec[1] = make_double4(o[1], o[2], l, 0.0);
}

// The mixed-precision (compact) bond record of bond_reinit from a bond
// vector o = r2 - r1, its length l = sqrt(o.o) and unit vector u = o (1/l)
// that the caller already holds (the same bits bond_reinit would compute).
template <class DevT, class GeoT>
__device__ __forceinline__ void bond_store(const DevT &d, const GeoT &g, int64_t b,
                                           const double *r1, const double *o, double l,
                                           const double *u) {
  double pos[3];
  for (int k = 0; k < 3; ++k) pos[k] = r1[k] + 0.5 * o[k];
  double sp[3];
  for (int k = 0; k < 3; ++k) {
    double si = 0.0;
    for (int j = 0; j < 3; ++j) si += g.hinv[3 * k + j] * pos[j];
    si -= (double)nint(si);
    sp[k] = si;
  }
  rec_write_compact(d.br_out, g, b, sp, u, l);
}

}  // namespace sc
