// sc_integrate_reg.hip -- register-resident integrator for filaments of up
// to 64 sites (the fast path of Filament::UpdatePosition,
// src/filament.cpp:394-428).
//
// A wavefront holds 64/G filaments; lane gl of a group owns site gl.  Every
// per-site value stays in registers: the +-1 and +-2 site neighbours come
// from lane shuffles (ds_bpermute, no LDS allocation), the two tridiagonal
// solves are parallel cyclic reduction and the metric determinant
// recursions (src/filament.cpp:613-623) are 2x2 matrix prefix products,
// both over in-group lane shuffles.  No LDS and no barriers: the
// wave runs its filaments start to finish without waiting on anything but
// its own loads.  The algebra uses the reduced forms of the fast path
// (mobility f/gperp + (1/gpar-1/gperp) t(t.f), bending (I - b b^T) a) and
// reciprocals by rcp_fast; the FP64 reference-order kernel remains the
// exact path.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_lanes.h"
#include "sc_math.h"

namespace sc {

using namespace lanes;

template <int G>
__global__ __launch_bounds__(256) void k_integrate_reg(Dev d, Geo g, const int32_t *list,
                                                       int count, int prec) {
  constexpr int NG = 64 / G;
  const Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int grp = lane / G, gl = lane % G;
  const int fidx = wave * NG + grp;
  const bool factive = fidx < count;
  const int f = factive ? list[fidx] : 0;
  const int64_t s0 = factive ? d.site_off[f] : 0;
  const int n = factive ? (int)(d.site_off[f + 1] - s0) : 0;
  const int64_t b0 = s0 - f;
  const int64_t ns = g.ns;
  const bool mid = ctl->midstep != 0;
  const int64_t step = ctl->cur_step;
  const int i = gl;
  const bool act = i < n;
  const int m = n - 1;  // bonds / tension unknowns
  double bl = 1.0, kappa = 0.0, gpar = 1.0, gperp = 1.0, sgpar = 0.0, sgperp = 0.0;
  if (factive) {
    const double *fp = d.fpar + 8 * f;
    bl = fp[0];
    kappa = fp[1];
    gpar = fp[2];
    gperp = fp[3];
    sgpar = fp[4];
    sgperp = fp[5];
  }
  const int64_t si = s0 + i;
  // ---- loads -------------------------------------------------------------
  double rp[3] = {0, 0, 0}, u[3] = {1.0, 0, 0}, Fb[3] = {0, 0, 0}, Tb[3] = {0, 0, 0};
  double fr[3] = {0, 0, 0}, acc = 0.0, dr0[3] = {0, 0, 0};
  {
    // bond orientation from the site positions in FP64 with the arithmetic
    // of bond_reinit (the record, compact on this path, keeps it rounded to
    // FP32 and is not read here); the last site takes bond m-1's
    double rc[3], o[3], ub[3], l = 0.0;
    for (int q = 0; q < 3; ++q) rc[q] = act ? d.r[q * ns + si] : 0.0;
    for (int q = 0; q < 3; ++q) {
      o[q] = d_next(rc[q]) - rc[q];
      l += o[q] * o[q];
    }
    const double il = rsq_fast(l);
    for (int q = 0; q < 3; ++q) ub[q] = o[q] * il;
    for (int q = 0; q < 3; ++q) {
      const double pv = d_prev(ub[q]);
      u[q] = i < m ? ub[q] : (i == m ? pv : (q == 0 ? 1.0 : 0.0));
    }
  }
  if (act) {
    for (int q = 0; q < 3; ++q) rp[q] = mid ? d.r[q * ns + si] : d.rp[q * ns + si];
    if (i < m) {
      for (int q = 0; q < 3; ++q) {
        Fb[q] = (double)d.bf32[q * g.nb + b0 + i];
        Tb[q] = (double)d.bt32[q * g.nb + b0 + i];
      }
    }
    if (!mid && g.stoch)
      for (int q = 0; q < 3; ++q) fr[q] = reinterpret_cast<const float *>(d.fr)[q * ns + si];
    if (!mid) {
      acc = d.acc[si];
      for (int q = 0; q < 3; ++q) dr0[q] = d.dr0[q * ns + si];
    }
  }
  // ---- ApplyInteractionForces (src/filament.cpp:992-1048) ---------------
  const double linv = rcp_fast(bl);
  double sf[3], pt[3];
  cross3(u, Tb, pt);
  for (int q = 0; q < 3; ++q) {
    pt[q] *= linv;
    sf[q] = 0.5 * Fb[q];
  }
  double sfm[3], ptm[3];
  v_prev(sf, sfm);
  v_prev(pt, ptm);
  double fs[3] = {0.0, 0.0, 0.0};
  if (i >= 1)
    for (int q = 0; q < 3; ++q) fs[q] = fs[q] + sfm[q] + (-ptm[q]);
  if (i <= n - 2)
    for (int q = 0; q < 3; ++q) fs[q] = fs[q] + sf[q] + pt[q];
  // ---- angles and tangents ----------------------------------------------
  double um[3], up[3];
  v_prev(u, um);
  v_next(u, up);
  const double cs = (i < n - 2) ? dot(u, up) : 0.0;
  double t[3];
  if (i == 0 || i >= n - 1) {
    for (int q = 0; q < 3; ++q) t[q] = u[q];
  } else {
    for (int q = 0; q < 3; ++q) t[q] = um[q] + u[q];
    const double inv = rsq_fast(dot(t, t));
    for (int q = 0; q < 3; ++q) t[q] *= inv;
  }
  // ---- random forces (midstep): construct and project ------------------
  if (mid) {
    if (g.stoch) {
      double xi[3] = {0, 0, 0};
      if (act) draw_xi(g, (int64_t)d.noise_off[f] + i, (step - 1) >> 1, xi);
      const double txi = dot(t, xi);
      for (int q = 0; q < 3; ++q) fr[q] = sgperp * xi[q] + (sgpar - sgperp) * t[q] * txi;
      double frp[3], frn[3];
      v_next(fr, frn);
      for (int q = 0; q < 3; ++q) frp[q] = frn[q] - fr[q];
      const double rhs = dot(frp, u);
      const double csm = d_prev(cs);
      // G = tridiag(-cos, 2, -cos) on m unknowns
      const double Tp = pcr<G>(i >= 1 ? -csm : 0.0, 2.0, i < m - 1 ? -cs : 0.0, rhs, m, lane, gl);
      const double Tm = d_prev(Tp);
      if (i == 0) {
        for (int q = 0; q < 3; ++q) fr[q] = fr[q] + Tp * u[q];
      } else if (i == n - 1) {
        for (int q = 0; q < 3; ++q) fr[q] = fr[q] - Tm * um[q];
      } else {
        for (int q = 0; q < 3; ++q) fr[q] = fr[q] + Tp * u[q] - Tm * um[q];
      }
      // stored in FP32 (both half steps use the rounded value)
      for (int q = 0; q < 3; ++q) fr[q] = (double)(float)fr[q];
      if (act)
        for (int q = 0; q < 3; ++q) reinterpret_cast<float *>(d.fr)[q * ns + si] = (float)fr[q];
    }
    if (act)
      for (int q = 0; q < 3; ++q) d.rp[q * ns + si] = rp[q];
  }
  if (g.stoch)
    for (int q = 0; q < 3; ++q) fs[q] += fr[q];

  // ---- bending: metric determinants as 2x2 prefix/suffix products -------
  {
    const double csm = d_prev(cs), csmm = d_prev(csm);
    // det_t: lane i >= 2 holds M_i = [[2, -c_{i-2}^2], [1, 0]]
    double P[4] = {1.0, 0.0, 0.0, 1.0};
    if (i >= 2 && i < n) {
      P[0] = 2.0;
      P[1] = -csmm * csmm;
      P[2] = 1.0;
      P[3] = 0.0;
    }
    mat_scan<G>(P, lane, gl, -1);
    const double det_t = (i == 0) ? 1.0 : P[0] * 2.0 + P[1] * 1.0;
    // det_b[j] at lane j: N_j = [[2, -c_{j-1}^2], [1, 0]] for 1 <= j <= n-2
    double S[4] = {1.0, 0.0, 0.0, 1.0};
    if (i >= 1 && i <= n - 2) {
      S[0] = 2.0;
      S[1] = -csm * csm;
      S[2] = 1.0;
      S[3] = 0.0;
    }
    mat_scan<G>(S, lane, gl, +1);
    const double det_b = S[0] * 2.0 + S[1] * 1.0;
    const double det_g = __shfl(det_t, lane - gl + (n > 0 ? n - 1 : 0), 64);
    double db3 = g_from<G>(det_b, lane, gl, 3, 1.0);
    if (i + 3 >= n) db3 = 1.0;  // det_b[n] = 1
    double ke = 0.0;
    if (i < n - 2) {
      const double ginv = cs * det_t * db3 * rcp_fast(det_g);
      ke = (kappa + bl * ginv) * (linv * linv);
    }
    const double kem = d_prev(ke), kemm = d_prev(kem);
    double umm[3];
    v_prev(um, umm);
    double f3[3] = {0.0, 0.0, 0.0}, pa[3], pb[3];
    if (i > 1 && act) {
      proj_perp(umm, um, pa);
      for (int q = 0; q < 3; ++q) f3[q] += kemm * pa[q];
    }
    if (i > 0 && i < n - 1) {
      proj_perp(u, um, pa);
      proj_perp(um, u, pb);
      for (int q = 0; q < 3; ++q) f3[q] += kem * (pa[q] - pb[q]);
    }
    if (i < n - 2) {
      proj_perp(up, u, pa);
      for (int q = 0; q < 3; ++q) f3[q] -= ke * pa[q];
    }
    for (int q = 0; q < 3; ++q) fs[q] += f3[q];
  }

  // ---- tensions (src/filament.cpp:778-838) -------------------------------
  const double igpar = rcp_fast(gpar), igperp = rcp_fast(gperp), dg = igpar - igperp;
  double T;
  {
    double tn[3], fn[3], m1[3], m2[3], fd[3];
    v_next(t, tn);
    v_next(fs, fn);
    mob(igperp, dg, t, fs, m1);
    mob(igperp, dg, tn, fn, m2);
    for (int q = 0; q < 3; ++q) fd[q] = m2[q] - m1[q];
    const double rhs = dot(u, fd);
    const double a1 = dot(t, u), a2 = dot(tn, u);
    const double diag = 2.0 * igperp + dg * (a1 * a1 + a2 * a2);
    // coupling of rows i-1 and i, computed at row i
    const double hu = -igperp * dot(u, um) - dg * (dot(t, um) * dot(t, u));
    // the shuffle runs on every lane: under a lane-dependent condition the
    // lanes it reads from could be switched off
    const double hun = d_next(hu);
    const double lower = (i >= 1) ? hu : 0.0;
    const double upper = (i < m - 1) ? hun : 0.0;
    T = pcr<G>(lower, diag, upper, rhs, m, lane, gl);
  }
  if (i < m) d.tension[b0 + i] = T;

  // ---- positions (src/filament.cpp:840-927) ------------------------------
  const double delta = mid ? 0.5 * g.delta : g.delta;
  const double Tm = d_prev(T);
  if (i == 0) {
    for (int q = 0; q < 3; ++q) fs[q] += T * u[q];
  } else if (i == n - 1) {
    for (int q = 0; q < 3; ++q) fs[q] += -Tm * um[q];
  } else {
    for (int q = 0; q < 3; ++q) fs[q] += T * u[q] - Tm * um[q];
  }
  double rn[3], mv[3];
  mob(igperp, dg, t, fs, mv);
  for (int q = 0; q < 3; ++q) rn[q] = rp[q] + delta * mv[q];
  double rnn[3], uo[3];
  v_next(rn, rnn);
  int bad = 0;
  {
    double rd[3];
    for (int q = 0; q < 3; ++q) rd[q] = rnn[q] - rn[q];
    const double dd = dot(rd, rd);
    const double inv = rsq_fast(dd);
    const double mm = dd * inv;
    for (int q = 0; q < 3; ++q) uo[q] = rd[q] * inv;
    if (i < m) {
      const double diff = bl - mm;
      if ((diff < 0 ? -diff : diff) * linv > 1e-3) bad = 1;
    }
  }
  const unsigned long long bal = __ballot(bad);
  const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << (grp * G));
  if ((bal & gmask) != 0) {
    // bond renormalisation: cumulative sums of l u over the chain
    int32_t sw = 1;
    if (factive) sw = d.norm[2 * f];
    sw = !sw;
    double step3[3];
    for (int q = 0; q < 3; ++q) step3[q] = (i < m) ? bl * uo[q] : 0.0;
    if (sw) {
      // tail to head: r_i = r_0 + sum_{k < i} l u_k
      const double r0x = __shfl(rn[0], lane - gl, 64), r0y = __shfl(rn[1], lane - gl, 64),
                   r0z = __shfl(rn[2], lane - gl, 64);
      const double r0[3] = {r0x, r0y, r0z};
      for (int q = 0; q < 3; ++q) {
        const double inc = sum_scan<G>(step3[q], lane, gl, -1) - step3[q];
        if (i >= 1) rn[q] = r0[q] + inc;
      }
    } else {
      // head to tail: r_i = r_{n-1} - sum_{k >= i} l u_k
      const int last = lane - gl + (n > 0 ? n - 1 : 0);
      const double rl[3] = {__shfl(rn[0], last, 64), __shfl(rn[1], last, 64),
                            __shfl(rn[2], last, 64)};
      for (int q = 0; q < 3; ++q) {
        const double suf = sum_scan<G>(step3[q], lane, gl, +1);
        if (i < n - 1) rn[q] = rl[q] - suf;
      }
    }
    v_next(rn, rnn);
    if (gl == 0 && factive) {
      d.norm[2 * f] = sw;
      d.norm[2 * f + 1] = 1;
    }
  } else if (gl == 0 && factive) {
    d.norm[2 * f + 1] += 1;
  }

  // ---- write back, bond records, rebuild bookkeeping ---------------------
  double lmax = 0.0;
  if (act) {
    for (int q = 0; q < 3; ++q) d.r[q * ns + si] = rn[q];
    if (i < m) bond_reinit(d, g, b0 + i, rn, rnn, prec);
    if (!mid) {
      d.acc_prev[si] = acc;
      for (int rep = 0; rep < 2; ++rep)
        for (int q = 0; q < 3; ++q) {
          const double dd = rn[q] - dr0[q];
          acc += dd * dd;
        }
      d.acc[si] = acc;
      lmax = acc;
    }
  }
  if (!mid) {
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(lmax, o, 64);
      lmax = y > lmax ? y : lmax;
    }
    if (lane == 0)
      atomicMax(&d.dr_slots[(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) &
                            (kDrSlots - 1)],
                (unsigned long long)__double_as_longlong(lmax));
  }
}

bool launch_integrate_reg(const Dev &d, const Geo &g, const int32_t *bin_list, int bin_count,
                          int bin_maxn, int precision, int waves, hipStream_t st) {
  if (precision == SC_PREC_FP64 || bin_maxn > 64) return false;
  if (bin_count <= 0) return true;
  const int per_block = (64 / bin_maxn) * waves;
  dim3 grid((bin_count + per_block - 1) / per_block), block(64 * waves);
  switch (bin_maxn) {
    case 8:
      hipLaunchKernelGGL(k_integrate_reg<8>, grid, block, 0, st, d, g, bin_list, bin_count, precision);
      break;
    case 16:
      hipLaunchKernelGGL(k_integrate_reg<16>, grid, block, 0, st, d, g, bin_list, bin_count, precision);
      break;
    case 32:
      hipLaunchKernelGGL(k_integrate_reg<32>, grid, block, 0, st, d, g, bin_list, bin_count, precision);
      break;
    default:
      hipLaunchKernelGGL(k_integrate_reg<64>, grid, block, 0, st, d, g, bin_list, bin_count, precision);
      break;
  }
  return true;
}

}  // namespace sc
