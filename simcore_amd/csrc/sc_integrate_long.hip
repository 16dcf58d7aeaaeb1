// sc_integrate_long.hip -- register-resident fast-path integrator with
// several sites per lane (Filament::UpdatePosition,
// src/filament.cpp:394-428): filaments of up to G*S sites, G lanes each,
// 64/G filaments per wavefront.
//
// Group lane l owns the S consecutive sites lS .. lS+S-1, so a neighbour is
// either in the lane's own registers or one lane away (a DPP move).
// Tridiagonal systems: log2(S) levels of cyclic reduction inside each lane
// leave one row per lane, solved by parallel cyclic reduction across the
// group's lanes, then back-substituted level by level.  Metric
// determinants: 2x2 matrix products, sequential inside a lane, scanned
// across lanes.  Renormalisation: prefix sums the same way.  No LDS, no
// barriers.  Against one site per lane (sc_integrate_reg.hip) this halves
// the lanes, hence the shuffles, per filament.  Same reduced algebra;
// reciprocals (lengths, pivots, determinants) by rcp_fast rather than IEEE
// division, bond records written by bond_store from the new lengths and
// unit vectors.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_lanes.h"
#include "sc_math.h"

namespace sc {

using namespace lanes;

namespace {

// o[k] = value at site gl*S + k + DELTA of the filament's lane segment
// (fb outside its Gf lanes); the shuffles run on every lane.  W is the
// width of the aligned lane group the segment lies in (it selects the lane
// moves), Gf <= W the segment's own width.
template <int W, int S, int DELTA>
__device__ __forceinline__ void site_at(const double (&v)[S], double (&o)[S], int lane, int gl,
                                        int Gf, double fb) {
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int t = k + DELTA;
    const int q = t >= 0 ? t / S : -((-t + S - 1) / S);
    const int ks = t - q * S;
    if (q == 0) {
      o[k] = v[ks];
    } else {
      const bool ok = gl + q >= 0 && gl + q < Gf;
      const double y = lane_shift_c<W>(v[ks], lane, q);
      o[k] = ok ? y : fb;
    }
  }
}

template <int W, int S, int DELTA>
__device__ __forceinline__ void site_at3(const double (&v)[S][3], double (&o)[S][3], int lane,
                                         int gl, int Gf) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double a[S], b[S];
#pragma unroll
    for (int k = 0; k < S; ++k) a[k] = v[k][q];
    site_at<W, S, DELTA>(a, b, lane, gl, Gf, 0.0);
#pragma unroll
    for (int k = 0; k < S; ++k) o[k][q] = b[k];
  }
}

// value of segment lane gl+delta; `fb` outside [0, Gf)
template <int W>
__device__ __forceinline__ double seg_from(double v, int lane, int gl, int Gf, int delta,
                                           double fb) {
  const int src = gl + delta;
  const bool ok = src >= 0 && src < Gf;
  const double x = lane_shift_c<W>(v, lane, delta);
  return ok ? x : fb;
}

// Parallel cyclic reduction over the segment's lanes (row gl in lane gl;
// rows >= m inert).  Levels beyond the segment's width leave every row
// unchanged (both couplings are zero by then), so W bounds the loop.  The
// values moved in from lanes outside the segment are not masked: they only
// enter multiplied by a coupling that is zero there (lo / hi), and every
// lane's rows are finite (other segments' rows, or inert ones), so they
// add zero.
template <int W>
__device__ __forceinline__ double pcr_seg(double a, double b, double c, double d, int m, int lane,
                                          int gl, int Gf) {
  (void)Gf;
  if (gl >= m) {
    a = 0.0;
    b = 1.0;
    c = 0.0;
    d = 0.0;
  }
#pragma unroll
  for (int s = 1; s < W; s <<= 1) {
    const double rb = rcp_fast(b);
    const double am = lane_shift_c<W>(a, lane, -s);
    const double cm = lane_shift_c<W>(c, lane, -s);
    const double dm = lane_shift_c<W>(d, lane, -s);
    const double rm = lane_shift_c<W>(rb, lane, -s);
    const double ap = lane_shift_c<W>(a, lane, s);
    const double cp = lane_shift_c<W>(c, lane, s);
    const double dp = lane_shift_c<W>(d, lane, s);
    const double rp = lane_shift_c<W>(rb, lane, s);
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

// inclusive prefix (dir -1, newest on the left) / suffix (dir +1) product
// of 2x2 matrices over the segment; identity outside it
template <int W>
__device__ __forceinline__ void mat_scan_seg(double *P, int lane, int gl, int Gf, int dir) {
#pragma unroll
  for (int s = 1; s < W; s <<= 1) {
    double Q[4];
    const double id[4] = {1.0, 0.0, 0.0, 1.0};
    for (int k = 0; k < 4; ++k) Q[k] = seg_from<W>(P[k], lane, gl, Gf, dir * s, id[k]);
    double R[4];
    R[0] = P[0] * Q[0] + P[1] * Q[2];
    R[1] = P[0] * Q[1] + P[1] * Q[3];
    R[2] = P[2] * Q[0] + P[3] * Q[2];
    R[3] = P[2] * Q[1] + P[3] * Q[3];
    for (int k = 0; k < 4; ++k) P[k] = R[k];
  }
}

// inclusive prefix (dir -1) / suffix (dir +1) sum over the segment
template <int W>
__device__ __forceinline__ double sum_scan_seg(double v, int lane, int gl, int Gf, int dir) {
#pragma unroll
  for (int s = 1; s < W; s <<= 1) v += seg_from<W>(v, lane, gl, Gf, dir * s, 0.0);
  return v;
}

// Solves rows i = lS + k < m (lower a: x[i-1], diagonal b, upper c:
// x[i+1], right-hand side d); rows >= m are inert.  Result in x.
template <int W, int S>
__device__ __forceinline__ void solve_blocked(double (&a)[S], double (&b)[S], double (&c)[S],
                                              double (&d)[S], int m, int lane, int gl, int Gf,
                                              double (&x)[S]) {
#pragma unroll
  for (int k = 0; k < S; ++k)
    if (gl * S + k >= m) {
      a[k] = 0.0;
      b[k] = 1.0;
      c[k] = 0.0;
      d[k] = 0.0;
    }
  // cyclic reduction inside the lane: at stride st the rows k = 0 mod 2st
  // absorb their neighbours k - st (the previous lane's S - st for k = 0)
  // and k + st; the rows k = st mod 2st keep their coefficients for the
  // back substitution
#pragma unroll
  for (int st = 1; st < S; st <<= 1) {
    // shuffles on every lane (a lane-dependent condition around them would
    // switch off the lanes they read from), then the boundary select
    const bool first = gl == 0;
    const double ya = lane_shift<W, -1>(a[S - st], lane), yb = lane_shift<W, -1>(b[S - st], lane);
    const double yc = lane_shift<W, -1>(c[S - st], lane), yd = lane_shift<W, -1>(d[S - st], lane);
    const double pa = first ? 0.0 : ya, pb = first ? 1.0 : yb;
    const double pc = first ? 0.0 : yc, pd = first ? 0.0 : yd;
#pragma unroll
    for (int k = 0; k < S; k += 2 * st) {
      const double la = k == 0 ? pa : a[k - st], lb = k == 0 ? pb : b[k - st];
      const double lc = k == 0 ? pc : c[k - st], ld = k == 0 ? pd : d[k - st];
      const double ua = a[k + st], ub = b[k + st], uc = c[k + st], ud = d[k + st];
      const double al = -a[k] * rcp_fast(lb), ga = -c[k] * rcp_fast(ub);
      a[k] = al * la;
      c[k] = ga * uc;
      b[k] = b[k] + al * lc + ga * ua;
      d[k] = d[k] + al * ld + ga * ud;
    }
  }
  // one row per lane left: parallel cyclic reduction across the lanes
  const int m_red = (m + S - 1) / S;
  x[0] = pcr_seg<W>(a[0], b[0], c[0], d[0], m_red, lane, gl, Gf);
#pragma unroll
  for (int st = S / 2; st >= 1; st >>= 1) {
    const bool last = gl == Gf - 1;
    const double nx = lane_shift<W, 1>(x[0], lane);
#pragma unroll
    for (int k = st; k < S; k += 2 * st) {
      const double xl = x[k - st];
      const double xu = k + st < S ? x[k + st] : (last ? 0.0 : nx);
      x[k] = (d[k] - a[k] * xl - c[k] * xu) * rcp_fast(b[k]);
    }
  }
}

__device__ __forceinline__ void mat_mul(const double *P, const double *Q, double *R) {
  R[0] = P[0] * Q[0] + P[1] * Q[2];
  R[1] = P[0] * Q[1] + P[1] * Q[3];
  R[2] = P[2] * Q[0] + P[3] * Q[2];
  R[3] = P[2] * Q[1] + P[3] * Q[3];
}

// Prefix (forward) products of per-site 2x2 matrices M (newest on the
// left): out[k] = M_{lS+k} ... M_0 applied to [2, 1], first component.
template <int W, int S>
__device__ __forceinline__ void det_forward(const double (&M)[S][4], double (&out)[S],
                                            int lane, int gl, int Gf) {
  double L[S][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) L[0][q] = M[0][q];
#pragma unroll
  for (int k = 1; k < S; ++k) mat_mul(M[k], L[k - 1], L[k]);
  double A[4] = {L[S - 1][0], L[S - 1][1], L[S - 1][2], L[S - 1][3]};
  mat_scan_seg<W>(A, lane, gl, Gf, -1);  // A_l ... A_0
  double E[4];
  const double id[4] = {1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double y = lane_shift<W, -1>(A[q], lane);
    E[q] = gl > 0 ? y : id[q];
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    double P[4];
    mat_mul(L[k], E, P);
    out[k] = P[0] * 2.0 + P[1] * 1.0;
  }
}

// Suffix products: out[k] = N_{lS+k} ... N_last applied to [2, 1].
template <int W, int S>
__device__ __forceinline__ void det_backward(const double (&N)[S][4], double (&out)[S],
                                             int lane, int gl, int Gf) {
  double R[S][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) R[S - 1][q] = N[S - 1][q];
#pragma unroll
  for (int k = S - 2; k >= 0; --k) mat_mul(N[k], R[k + 1], R[k]);
  double B[4] = {R[0][0], R[0][1], R[0][2], R[0][3]};
  mat_scan_seg<W>(B, lane, gl, Gf, +1);  // B_l ... B_last
  double F[4];
  const double id[4] = {1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double y = lane_shift<W, 1>(B[q], lane);
    F[q] = gl < Gf - 1 ? y : id[q];
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    double P[4];
    mat_mul(R[k], F, P);
    out[k] = P[0] * 2.0 + P[1] * 1.0;
  }
}

}  // namespace

// One filament's lane segment: lanes [lane - gl, lane - gl + Gf) of the
// wavefront hold its sites, two (S) per lane; the segment lies inside an
// aligned group of W lanes (16, 32 or 64), which selects the lane moves.
// Lanes with `factive` false take part in every lane move and write
// nothing.  Everything stays in registers; there is no LDS and no barrier.
// MID: the step is a midstep (compiled separately from the full step, so
// each gets its own register allocation; the midstep's random-force solve
// sets the larger one).  `slot`: the displacement-maximum slot of the wave.
template <int W, int S, bool MID>
__device__ __forceinline__ void long_body(const Dev &d, const Geo &g, bool factive, int f,
                                          int gl, int Gf, int slot) {
  const Ctl *ctl = d.ctl;
  const int lane = threadIdx.x & 63;
  const int64_t s0 = factive ? d.site_off[f] : 0;
  const int n = factive ? (int)(d.site_off[f + 1] - s0) : 0;
  const int64_t b0 = s0 - f;
  const int64_t ns = g.ns;
  constexpr bool mid = MID;
  const int64_t step = ctl->cur_step;
  const int m = n - 1;
  double bl = 1.0, kappa = 0.0, gpar = 1.0, gperp = 1.0, sgpar = 0.0, sgperp = 0.0;
  int64_t noise0 = 0;
  if (factive) {
    const double *fp = d.fpar + 8 * f;
    bl = fp[0];
    kappa = fp[1];
    gpar = fp[2];
    gperp = fp[3];
    sgpar = fp[4];
    sgperp = fp[5];
    noise0 = d.noise_off[f];
  }
  const int base = gl * S;

  // ---- loads (values needed only at the end are loaded there, which keeps
  // register pressure, hence occupancy, in check) --------------------------
  double u[S][3], Fb[S][3], Tb[S][3], fr[S][3];
  {
    // bond orientations from the site positions in FP64 with the arithmetic
    // of bond_reinit (the record, compact on this path, keeps them rounded
    // to FP32 and is not read here); the last site takes bond m-1's
    double rc[S][3], rnx[S][3], ub[S][3], ubm[S][3];
    // an opaque copy of the pointer: the midstep's second load of r (at the
    // end) must not be merged with this one, which would keep the positions
    // in registers across the whole kernel
    const double *rsrc = d.r;
    asm volatile("" : "+s"(rsrc));
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      for (int q = 0; q < 3; ++q) rc[k][q] = i < n ? rsrc[q * ns + s0 + i] : 0.0;
    }
    site_at3<W, S, 1>(rc, rnx, lane, gl, Gf);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      double o[3], l = 0.0;
      for (int q = 0; q < 3; ++q) {
        o[q] = rnx[k][q] - rc[k][q];
        l += o[q] * o[q];
      }
      const double il = rsq_fast(l);
      for (int q = 0; q < 3; ++q) ub[k][q] = o[q] * il;
    }
    site_at3<W, S, -1>(ub, ubm, lane, gl, Gf);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      for (int q = 0; q < 3; ++q)
        u[k][q] = i < m ? ub[k][q] : (i == m ? ubm[k][q] : (q == 0 ? 1.0 : 0.0));
    }
  }
#pragma unroll
  for (int k = 0; k < S; ++k)
    for (int q = 0; q < 3; ++q) fr[k][q] = 0.0;

  const double linv = rcp_fast(bl);

  // ---- angles and tangents ----------------------------------------------
  double um[S][3], up[S][3], t[S][3], cs[S];
  site_at3<W, S, -1>(u, um, lane, gl, Gf);
  site_at3<W, S, 1>(u, up, lane, gl, Gf);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = base + k;
    cs[k] = i < n - 2 ? dot(u[k], up[k]) : 0.0;
    if (i == 0 || i >= n - 1) {
      for (int q = 0; q < 3; ++q) t[k][q] = u[k][q];
    } else {
      for (int q = 0; q < 3; ++q) t[k][q] = um[k][q] + u[k][q];
      const double inv = rsq_fast(dot(t[k], t[k]));
      for (int q = 0; q < 3; ++q) t[k][q] *= inv;
    }
  }
  double csm[S];
  site_at<W, S, -1>(cs, csm, lane, gl, Gf, 0.0);

  // ---- random forces (midstep) ------------------------------------------
  if (mid) {
    if (g.stoch) {
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = base + k;
        double xi[3] = {0.0, 0.0, 0.0};
        if (i < n) draw_xi(g, noise0 + i, (step - 1) >> 1, xi);
        const double txi = dot(t[k], xi);
        for (int q = 0; q < 3; ++q)
          fr[k][q] = sgperp * xi[q] + (sgpar - sgperp) * t[k][q] * txi;
      }
      double frn[S][3], a[S], b[S], c[S], rh[S], Tp[S], Tm[S];
      site_at3<W, S, 1>(fr, frn, lane, gl, Gf);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = base + k;
        double df[3];
        for (int q = 0; q < 3; ++q) df[q] = frn[k][q] - fr[k][q];
        rh[k] = dot(df, u[k]);
        a[k] = i >= 1 ? -csm[k] : 0.0;
        b[k] = 2.0;
        c[k] = i < m - 1 ? -cs[k] : 0.0;
      }
      solve_blocked<W, S>(a, b, c, rh, m, lane, gl, Gf, Tp);
      site_at<W, S, -1>(Tp, Tm, lane, gl, Gf, 0.0);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = base + k;
        if (i == 0) {
          for (int q = 0; q < 3; ++q) fr[k][q] = fr[k][q] + Tp[k] * u[k][q];
        } else if (i == n - 1) {
          for (int q = 0; q < 3; ++q) fr[k][q] = fr[k][q] - Tm[k] * um[k][q];
        } else {
          for (int q = 0; q < 3; ++q) fr[k][q] = fr[k][q] + Tp[k] * u[k][q] - Tm[k] * um[k][q];
        }
        // stored in FP32 (both half steps use the rounded value)
        for (int q = 0; q < 3; ++q) fr[k][q] = (double)(float)fr[k][q];
        if (i < n)
          for (int q = 0; q < 3; ++q)
            reinterpret_cast<float *>(d.fr)[q * ns + s0 + i] = (float)fr[k][q];
      }
    }
  }
  // ---- ApplyInteractionForces (src/filament.cpp:992-1048) ---------------
  // (after the random forces: the bond sums are loaded here, so they and
  // the site forces are not live across the random-force solve)
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = base + k;
    for (int q = 0; q < 3; ++q) {
      Fb[k][q] = 0.0;
      Tb[k][q] = 0.0;
    }
    if (i < m)
      for (int q = 0; q < 3; ++q) {
        Fb[k][q] = (double)d.bf32[q * g.nb + b0 + i];
        Tb[k][q] = (double)d.bt32[q * g.nb + b0 + i];
      }
  }
  double sf[S][3], pt[S][3], sfm[S][3], ptm[S][3], fs[S][3];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    cross3(u[k], Tb[k], pt[k]);
    for (int q = 0; q < 3; ++q) {
      pt[k][q] *= linv;
      sf[k][q] = 0.5 * Fb[k][q];
    }
  }
  site_at3<W, S, -1>(sf, sfm, lane, gl, Gf);
  site_at3<W, S, -1>(pt, ptm, lane, gl, Gf);
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = base + k;
    for (int q = 0; q < 3; ++q) fs[k][q] = 0.0;
    if (i >= 1)
      for (int q = 0; q < 3; ++q) fs[k][q] = fs[k][q] + sfm[k][q] + (-ptm[k][q]);
    if (i <= n - 2)
      for (int q = 0; q < 3; ++q) fs[k][q] = fs[k][q] + sf[k][q] + pt[k][q];
  }

  if (g.stoch)
#pragma unroll
    for (int k = 0; k < S; ++k) {
      // full step: the projected random force of the midstep
      if (!mid && base + k < n)
        for (int q = 0; q < 3; ++q)
          fr[k][q] = reinterpret_cast<const float *>(d.fr)[q * ns + s0 + base + k];
      for (int q = 0; q < 3; ++q) fs[k][q] += fr[k][q];
    }

  // ---- bending (src/filament.cpp:608-776), reduced form ------------------
  {
    double csmm[S], M[S][4], N[S][4], det_t[S], det_b[S], db3[S], ke[S];
    site_at<W, S, -1>(csm, csmm, lane, gl, Gf, 0.0);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      const bool sm = i >= 2 && i < n;
      M[k][0] = sm ? 2.0 : 1.0;
      M[k][1] = sm ? -csmm[k] * csmm[k] : 0.0;
      M[k][2] = sm ? 1.0 : 0.0;
      M[k][3] = sm ? 0.0 : 1.0;
    }
    det_forward<W, S>(M, det_t, lane, gl, Gf);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      const bool sn = i >= 1 && i <= n - 2;
      N[k][0] = sn ? 2.0 : 1.0;
      N[k][1] = sn ? -csm[k] * csm[k] : 0.0;
      N[k][2] = sn ? 1.0 : 0.0;
      N[k][3] = sn ? 0.0 : 1.0;
    }
    det_backward<W, S>(N, det_b, lane, gl, Gf);
    // det_t at the last site, det_b three sites ahead (det_b[n] = 1)
    const int last = n > 0 ? n - 1 : 0;
    double pick = 0.0;
#pragma unroll
    for (int k = 0; k < S; ++k)
      if (k == last % S) pick = det_t[k];
    const double det_g = __shfl(pick, lane - gl + last / S, 64);
    site_at<W, S, 3>(det_b, db3, lane, gl, Gf, 1.0);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      if (i == 0) det_t[k] = 1.0;
      if (i + 3 >= n) db3[k] = 1.0;
      ke[k] = 0.0;
      if (i < n - 2) {
        const double ginv = cs[k] * det_t[k] * db3[k] * rcp_fast(det_g);
        ke[k] = (kappa + bl * ginv) * (linv * linv);
      }
    }
    double kem[S], kemm[S], umm[S][3];
    site_at<W, S, -1>(ke, kem, lane, gl, Gf, 0.0);
    site_at<W, S, -2>(ke, kemm, lane, gl, Gf, 0.0);
    site_at3<W, S, -2>(u, umm, lane, gl, Gf);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      double f3[3] = {0.0, 0.0, 0.0}, pa[3], pb[3];
      if (i > 1 && i < n) {
        proj_perp(umm[k], um[k], pa);
        for (int q = 0; q < 3; ++q) f3[q] += kemm[k] * pa[q];
      }
      if (i > 0 && i < n - 1) {
        proj_perp(u[k], um[k], pa);
        proj_perp(um[k], u[k], pb);
        for (int q = 0; q < 3; ++q) f3[q] += kem[k] * (pa[q] - pb[q]);
      }
      if (i < n - 2) {
        proj_perp(up[k], u[k], pa);
        for (int q = 0; q < 3; ++q) f3[q] -= ke[k] * pa[q];
      }
      for (int q = 0; q < 3; ++q) fs[k][q] += f3[q];
    }
  }

  // ---- tensions (src/filament.cpp:778-838) -------------------------------
  const double igpar = rcp_fast(gpar), igperp = rcp_fast(gperp), dg = igpar - igperp;
  double T[S], Tm[S];
  {
    double tn[S][3], fn[S][3], hu[S], hun[S], a[S], b[S], c[S], rh[S];
    site_at3<W, S, 1>(t, tn, lane, gl, Gf);
    site_at3<W, S, 1>(fs, fn, lane, gl, Gf);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      double m1[3], m2[3], fd[3];
      mob(igperp, dg, t[k], fs[k], m1);
      mob(igperp, dg, tn[k], fn[k], m2);
      for (int q = 0; q < 3; ++q) fd[q] = m2[q] - m1[q];
      rh[k] = dot(u[k], fd);
      const double a1 = dot(t[k], u[k]), a2 = dot(tn[k], u[k]);
      b[k] = 2.0 * igperp + dg * (a1 * a1 + a2 * a2);
      hu[k] = -igperp * dot(u[k], um[k]) - dg * (dot(t[k], um[k]) * dot(t[k], u[k]));
    }
    site_at<W, S, 1>(hu, hun, lane, gl, Gf, 0.0);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = base + k;
      a[k] = i >= 1 ? hu[k] : 0.0;
      c[k] = i < m - 1 ? hun[k] : 0.0;
    }
    solve_blocked<W, S>(a, b, c, rh, m, lane, gl, Gf, T);
  }
  site_at<W, S, -1>(T, Tm, lane, gl, Gf, 0.0);
#pragma unroll
  for (int k = 0; k < S; ++k)
    if (base + k < m) d.tension[b0 + base + k] = T[k];

  // ---- positions (src/filament.cpp:840-927) ------------------------------
  // midstep start r_prev: the current positions on a midstep (stored as the
  // new r_prev), the stored ones on a full step
  const double delta = mid ? 0.5 * g.delta : g.delta;
  double rn[S][3], rnn[S][3], uo[S][3];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = base + k;
    double rp[3] = {0.0, 0.0, 0.0};
    if (i < n) {
      const int64_t si = s0 + i;
      for (int q = 0; q < 3; ++q) rp[q] = mid ? d.r[q * ns + si] : d.rp[q * ns + si];
      if (mid)
        for (int q = 0; q < 3; ++q) d.rp[q * ns + si] = rp[q];
    }
    if (i == 0) {
      for (int q = 0; q < 3; ++q) fs[k][q] += T[k] * u[k][q];
    } else if (i == n - 1) {
      for (int q = 0; q < 3; ++q) fs[k][q] += -Tm[k] * um[k][q];
    } else {
      for (int q = 0; q < 3; ++q) fs[k][q] += T[k] * u[k][q] - Tm[k] * um[k][q];
    }
    double mv[3];
    mob(igperp, dg, t[k], fs[k], mv);
    for (int q = 0; q < 3; ++q) rn[k][q] = rp[q] + delta * mv[q];
  }
  // new bond lengths mm and unit vectors uo, kept for the bond records
  double mm[S];
  site_at3<W, S, 1>(rn, rnn, lane, gl, Gf);
  int bad = 0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = base + k;
    double rd[3];
    for (int q = 0; q < 3; ++q) rd[q] = rnn[k][q] - rn[k][q];
    const double dd = dot(rd, rd);
    const double inv = rsq_fast(dd);
    mm[k] = dd * inv;
    for (int q = 0; q < 3; ++q) uo[k][q] = rd[q] * inv;
    if (i < m) {
      const double diff = bl - mm[k];
      if ((diff < 0 ? -diff : diff) * linv > 1e-3) bad = 1;
    }
  }
  const unsigned long long bal = __ballot(bad);
  const unsigned long long gmask =
      (Gf >= 64) ? ~0ull : (((1ull << Gf) - 1ull) << (lane - gl));
  const bool need = (bal & gmask) != 0;
  if (bal != 0) {
    // bond renormalisation (src/filament.cpp:904-924): cumulative sums of
    // l u along the chain, tail to head or head to tail by the switch.  The
    // branch is taken by the whole wavefront and both directions are summed
    // on every lane: the segments of a wavefront share lane groups, whose
    // wide moves pass values through other segments' lanes, so every lane
    // must take part; each segment keeps the result it needs.
    int32_t sw = factive ? d.norm[2 * f] : 1;
    sw = !sw;
    double st3[S][3];
#pragma unroll
    for (int k = 0; k < S; ++k)
      for (int q = 0; q < 3; ++q) st3[k][q] = base + k < m ? bl * uo[k][q] : 0.0;
    const int last = n > 0 ? n - 1 : 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      // tail to head: r_i = r_0 + sum_{j < i} l u_j
      double locp[S], run = 0.0;
#pragma unroll
      for (int k = 0; k < S; ++k) {
        locp[k] = run;  // exclusive inside the lane
        run += st3[k][q];
      }
      const double inclp = sum_scan_seg<W>(run, lane, gl, Gf, -1);
      const double yp = lane_shift<W, -1>(inclp, lane);
      const double before = gl > 0 ? yp : 0.0;
      const double r0 = __shfl(rn[0][q], lane - gl, 64);
      // head to tail: r_i = r_{n-1} - sum_{j >= i} l u_j
      double locs[S];
      run = 0.0;
#pragma unroll
      for (int k = S - 1; k >= 0; --k) {
        run += st3[k][q];
        locs[k] = run;  // inclusive suffix inside the lane
      }
      const double incls = sum_scan_seg<W>(run, lane, gl, Gf, +1);
      const double ys = lane_shift<W, 1>(incls, lane);
      const double after = gl < Gf - 1 ? ys : 0.0;
      double pick = 0.0;
#pragma unroll
      for (int k = 0; k < S; ++k)
        if (k == last % S) pick = rn[k][q];
      const double rl = __shfl(pick, lane - gl + last / S, 64);
#pragma unroll
      for (int k = 0; k < S; ++k) {
        if (!need) continue;
        if (sw) {
          if (base + k >= 1) rn[k][q] = r0 + (before + locp[k]);
        } else {
          if (base + k < n - 1) rn[k][q] = rl - (after + locs[k]);
        }
      }
    }
    site_at3<W, S, 1>(rn, rnn, lane, gl, Gf);
#pragma unroll
    for (int k = 0; k < S; ++k) {
      double rd[3];
      for (int q = 0; q < 3; ++q) rd[q] = rnn[k][q] - rn[k][q];
      const double dd = dot(rd, rd);
      const double inv = rsq_fast(dd);
      mm[k] = dd * inv;
      for (int q = 0; q < 3; ++q) uo[k][q] = rd[q] * inv;
    }
    if (need && gl == 0 && factive) {
      d.norm[2 * f] = sw;
      d.norm[2 * f + 1] = 1;
    }
  }
  if (!need && gl == 0 && factive) d.norm[2 * f + 1] += 1;

  // ---- write back, bond records, rebuild bookkeeping ---------------------
  double lmax = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = base + k;
    if (i >= n) continue;
    const int64_t si = s0 + i;
    for (int q = 0; q < 3; ++q) d.r[q * ns + si] = rn[k][q];
    if (i < m) {
      // the record bond_reinit (mixed) computes, from the values at hand
      double rd[3];
      for (int q = 0; q < 3; ++q) rd[q] = rnn[k][q] - rn[k][q];
      bond_store(d, g, b0 + i, rn[k], rd, mm[k], uo[k]);
    }
    if (!mid) {
      double a = d.acc[si], dr0[3];
      for (int q = 0; q < 3; ++q) dr0[q] = d.dr0[q * ns + si];
      d.acc_prev[si] = a;
      for (int rep = 0; rep < 2; ++rep)
        for (int q = 0; q < 3; ++q) {
          const double dd = rn[k][q] - dr0[q];
          a += dd * dd;
        }
      d.acc[si] = a;
      lmax = a > lmax ? a : lmax;
    }
  }
  if (!mid) {
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(lmax, o, 64);
      lmax = y > lmax ? y : lmax;
    }
    if (lane == 0)
      atomicMax(&d.dr_slots[slot & (kDrSlots - 1)],
                (unsigned long long)__double_as_longlong(lmax));
  }
}

// two sites per lane: at most 128 VGPRs, four waves per SIMD (the midstep
// needs 129 left to itself); four sites per lane: two waves.  Length-binned
// layout: G lanes per filament, 64/G filaments per wave, wave `wave` of a
// bin holding list entries wave*64/G ... (wave+1)*64/G - 1.
template <int G, int S, bool MID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S == 2 ? 4 : 1)))
void k_integrate_long(Dev d, Geo g, const int32_t *list,
                                                        int count) {
  Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  if ((ctl->midstep != 0) != MID) {  // launched for the wrong kind of step
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->err = SC_ERR_STATE;
      ctl->abort = 1;
    }
    return;
  }
  constexpr int NG = 64 / G;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int fidx = wave * NG + lane / G;
  const bool factive = fidx < count;
  long_body<G, S, MID>(d, g, factive, factive ? list[fidx] : 0, lane % G, G, wave);
}

// Packed layout: filaments of up to 128 sites placed as lane segments of
// their own width ceil(n/2) (lane table entry: filament << 8 | W code << 6
// | lane in segment, -1 for an idle lane), a segment of at most 16 lanes
// inside one 16-lane row, of 17-32 lanes inside one half, of 33-64 lanes
// from lane 0.  The waves are grouped by the lane-group width W of their
// widest segment (16, 32 or 64); one launch per width carries only its own
// lane moves and register budget.  Same arithmetic as the binned layout,
// far fewer idle lanes.
template <int W, bool MID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W == 64 ? 3 : 4)))
void k_integrate_packed(Dev d, Geo g, const int32_t *lane_tab, int n_waves) {
  Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  if ((ctl->midstep != 0) != MID) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->err = SC_ERR_STATE;
      ctl->abort = 1;
    }
    return;
  }
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= n_waves) return;
  const int lane = threadIdx.x & 63;
  const int32_t e = lane_tab[(int64_t)wave * 64 + lane];
  const bool factive = e >= 0;
  const int f = factive ? (e >> 8) : 0;
  const int gl = factive ? (e & 63) : 0;
  int Gf = 1;
  if (factive) Gf = (d.site_off[f + 1] - d.site_off[f] + 1) >> 1;
  long_body<W, 2, MID>(d, g, factive, f, gl, Gf, wave);
}

// All packed waves in one launch: the wave's lane-group width follows from
// its index (the waves are grouped by width, [0, cs1) 16, [cs1, cs2) 32,
// the rest 64) and selects the variant, uniformly per wave.  One register
// budget for the three variants (about 140 VGPRs, three waves per SIMD;
// the spills of a 128-register budget cost more than the fourth wave
// gains); fewer launches than one per width.
template <bool MID>
__device__ __forceinline__ void packed_any_body(const Dev &d, const Geo &g,
                                                const int32_t *lane_tab, int n_waves, int cs1,
                                                int cs2, int bid) {
  Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  if ((ctl->midstep != 0) != MID) {
    if (bid == 0 && threadIdx.x == 0) {
      ctl->err = SC_ERR_STATE;
      ctl->abort = 1;
    }
    return;
  }
  const int wave = bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= n_waves) return;
  const int lane = threadIdx.x & 63;
  const int32_t e = lane_tab[(int64_t)wave * 64 + lane];
  const bool factive = e >= 0;
  const int f = factive ? (e >> 8) : 0;
  const int gl = factive ? (e & 63) : 0;
  int Gf = 1;
  if (factive) Gf = (d.site_off[f + 1] - d.site_off[f] + 1) >> 1;
  if (wave < cs1)
    long_body<16, 2, MID>(d, g, factive, f, gl, Gf, wave);
  else if (wave < cs2)
    long_body<32, 2, MID>(d, g, factive, f, gl, Gf, wave);
  else
    long_body<64, 2, MID>(d, g, factive, f, gl, Gf, wave);
}

template <bool MID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_integrate_packed_any(Dev d, Geo g, const int32_t *lane_tab, int n_waves, int cs1,
                            int cs2) {
  packed_any_body<MID>(d, g, lane_tab, n_waves, cs1, cs2, blockIdx.x);
}

// Every filament of the step in one launch: the n_long filaments of 129-256
// sites (four sites per lane, one filament per wave) take the first waves,
// so these latency-bound waves start first and run beside the packed waves
// that follow, instead of in a launch of their own after them.  The same
// arithmetic as the two separate launches, bit for bit.
template <bool MID>
__device__ __forceinline__ void all_body(const Dev &d, const Geo &g, const int32_t *lane_tab,
                                         int n_waves, int cs1, int cs2,
                                         const int32_t *long_list, int n_long, int bid) {
  const int wave = bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave < n_long) {
    Ctl *ctl = d.ctl;
    if (ctl->abort) return;
    if ((ctl->midstep != 0) != MID) {
      if (wave == 0 && (threadIdx.x & 63) == 0) {
        ctl->err = SC_ERR_STATE;
        ctl->abort = 1;
      }
      return;
    }
    long_body<64, 4, MID>(d, g, true, long_list[wave], threadIdx.x & 63, 64, wave);
    return;
  }
  // the packed waves, numbered from 0 after the long ones (the block index
  // argument counts whole blocks, so the offset goes into the wave index)
  const int pw = wave - n_long;
  Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  if ((ctl->midstep != 0) != MID) {
    if (wave == 0 && (threadIdx.x & 63) == 0) {
      ctl->err = SC_ERR_STATE;
      ctl->abort = 1;
    }
    return;
  }
  if (pw >= n_waves) return;
  const int lane = threadIdx.x & 63;
  const int32_t e = lane_tab[(int64_t)pw * 64 + lane];
  const bool factive = e >= 0;
  const int f = factive ? (e >> 8) : 0;
  const int gl = factive ? (e & 63) : 0;
  int Gf = 1;
  if (factive) Gf = (d.site_off[f + 1] - d.site_off[f] + 1) >> 1;
  if (pw < cs1)
    long_body<16, 2, MID>(d, g, factive, f, gl, Gf, pw);
  else if (pw < cs2)
    long_body<32, 2, MID>(d, g, factive, f, gl, Gf, pw);
  else
    long_body<64, 2, MID>(d, g, factive, f, gl, Gf, pw);
}

template <bool MID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_integrate_all(Dev d, Geo g, const int32_t *lane_tab, int n_waves, int cs1, int cs2,
                     const int32_t *long_list, int n_long) {
  all_body<MID>(d, g, lane_tab, n_waves, cs1, cs2, long_list, n_long, blockIdx.x);
}

// Ensemble batches: every member's filaments as in its one-launch layout
// (those of 129-256 sites first, then the packed waves), one wave per
// block.  Odd steps (midsteps) use the members' odd-step pointers.
template <bool MID>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void k_group_integrate(const Member *__restrict__ ms, GroupGrid gg) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  all_body<MID>(M.d[MID ? 0 : 1], M.g, M.pack_tab, M.pack_waves, M.cs1, M.cs2, M.long_list,
                M.n_long, blockIdx.x - b0);
}

void launch_group_integrate(const Member *dm, const Member *hm, int n, bool odd,
                            hipStream_t st) {
  GroupGrid gg;
  int32_t t = 0;
  for (int m = 0; m < n; ++m) {
    gg.off[m] = t;
    t += hm[m].pack_waves + hm[m].n_long;
  }
  for (int m = n; m <= kMaxGroup; ++m) gg.off[m] = t;
  if (t <= 0) return;
  if (odd)
    hipLaunchKernelGGL((k_group_integrate<true>), dim3((unsigned)t), dim3(64), 0, st, dm, gg);
  else
    hipLaunchKernelGGL((k_group_integrate<false>), dim3((unsigned)t), dim3(64), 0, st, dm, gg);
}

void launch_integrate_packed_all(const Dev &d, const Geo &g, const int32_t *lane_tab,
                                 int n_waves, int cs1, int cs2, int waves, bool mid,
                                 hipStream_t st) {
  if (n_waves <= 0) return;
  dim3 grid((n_waves + waves - 1) / waves), block(64 * waves);
  if (mid)
    hipLaunchKernelGGL((k_integrate_packed_any<true>), grid, block, 0, st, d, g, lane_tab,
                       n_waves, cs1, cs2);
  else
    hipLaunchKernelGGL((k_integrate_packed_any<false>), grid, block, 0, st, d, g, lane_tab,
                       n_waves, cs1, cs2);
}

void launch_integrate_all(const Dev &d, const Geo &g, const int32_t *lane_tab, int n_waves,
                          int cs1, int cs2, const int32_t *long_list, int n_long, int waves,
                          bool mid, hipStream_t st) {
  const int total = n_waves + n_long;
  if (total <= 0) return;
  dim3 grid((total + waves - 1) / waves), block(64 * waves);
  if (mid)
    hipLaunchKernelGGL((k_integrate_all<true>), grid, block, 0, st, d, g, lane_tab, n_waves,
                       cs1, cs2, long_list, n_long);
  else
    hipLaunchKernelGGL((k_integrate_all<false>), grid, block, 0, st, d, g, lane_tab, n_waves,
                       cs1, cs2, long_list, n_long);
}

void launch_integrate_packed(const Dev &d, const Geo &g, const int32_t *lane_tab, int n_waves,
                             int width, int waves, bool mid, hipStream_t st) {
  if (n_waves <= 0) return;
  dim3 grid((n_waves + waves - 1) / waves), block(64 * waves);
#define SC_PACKED(WW)                                                                          \
  if (mid)                                                                                     \
    hipLaunchKernelGGL((k_integrate_packed<WW, true>), grid, block, 0, st, d, g, lane_tab,    \
                       n_waves);                                                               \
  else                                                                                         \
    hipLaunchKernelGGL((k_integrate_packed<WW, false>), grid, block, 0, st, d, g, lane_tab,   \
                       n_waves)
  if (width == 16) {
    SC_PACKED(16);
  } else if (width == 32) {
    SC_PACKED(32);
  } else {
    SC_PACKED(64);
  }
#undef SC_PACKED
}

bool launch_integrate_long(const Dev &d, const Geo &g, const int32_t *bin_list, int bin_count,
                           int bin_maxn, int precision, int waves, int layout, bool mid,
                           hipStream_t st) {
  if (precision == SC_PREC_FP64 || bin_maxn > 256) return false;
  // up to 64 sites: only in the two-sites-per-lane layout
  if (bin_maxn <= 64 && layout != 1) return false;
  if (bin_count <= 0) return true;
  const int G = bin_maxn <= 128 ? bin_maxn / 2 : 64;
  const int per_block = (64 / G) * waves;
  dim3 grid((bin_count + per_block - 1) / per_block), block(64 * waves);
#define SC_LONG(GG, SS)                                                                   \
  if (mid)                                                                                \
    hipLaunchKernelGGL((k_integrate_long<GG, SS, true>), grid, block, 0, st, d, g,       \
                       bin_list, bin_count);                                              \
  else                                                                                    \
    hipLaunchKernelGGL((k_integrate_long<GG, SS, false>), grid, block, 0, st, d, g,      \
                       bin_list, bin_count)
  switch (bin_maxn) {
    case 8: SC_LONG(4, 2); break;
    case 16: SC_LONG(8, 2); break;
    case 32: SC_LONG(16, 2); break;
    case 64: SC_LONG(32, 2); break;
    case 128: SC_LONG(64, 2); break;
    default: SC_LONG(64, 4); break;
  }
#undef SC_LONG
  return true;
}

}  // namespace sc
