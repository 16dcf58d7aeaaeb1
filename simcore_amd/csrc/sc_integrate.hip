// sc_integrate.hip -- Filament::UpdatePosition (src/filament.cpp:394-428)
// for all filaments of one length bin.
//
// Mapping: a wavefront holds 64/G filaments, G lanes each; lane gl of a
// group owns sites gl, gl+G, ..., gl+(S-1)G, so a bin serves filaments of up
// to G*S sites.  Per-site values live in LDS, one slice per group.
//
// The reference's sequential pieces -- the two Thomas solves
// (src/linear_algebra.cpp:164-197), the metric-determinant recursions
// (src/filament.cpp:613-623) and bond renormalisation (:904-924) -- run on
// each group's first lane with the recursion carried in registers, so the
// groups of a wave advance their chains together.  In FAST mode the
// tridiagonal solves instead use parallel cyclic reduction over the
// group's lanes (log2 n rounds); EXACT mode keeps Thomas, operation for
// operation, and reproduces the CPU path bit for bit.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

namespace {

// Thomas on m unknowns with the reference's operation order; lower[i]
// couples row i+1 to x[i], upper[i] couples row i to x[i+1].
__device__ __forceinline__ void thomas_reg(int m, const double *__restrict__ lo,
                                           const double *__restrict__ di,
                                           double *__restrict__ up,
                                           double *__restrict__ rh) {
  const int nn = m - 1;
  double cp = up[0] / di[0];
  double dp = rh[0] / di[0];
  up[0] = cp;
  rh[0] = dp;
  for (int i = 1; i < nn; ++i) {
    const double a = lo[i - 1];
    const double den = di[i] - a * cp;
    cp = up[i] / den;
    dp = (rh[i] - a * dp) / den;
    up[i] = cp;
    rh[i] = dp;
  }
  const double a = lo[nn - 1];
  double x = (rh[nn] - a * dp) / (di[nn] - a * cp);
  rh[nn] = x;
  for (int i = nn; i-- > 0;) {
    x = rh[i] - up[i] * x;
    rh[i] = x;
  }
}

// (I - b b^T) a = a - b (a.b): the metric-corrected bending term in three
// multiply-adds per component (FAST path).
__device__ __forceinline__ void proj_perp(const double *a, const double *b, double *out) {
  const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  for (int q = 0; q < 3; ++q) out[q] = a[q] - b[q] * ab;
}

// Site mobility applied to f: f/gperp + (1/gpar - 1/gperp) t (t.f).
__device__ __forceinline__ void mobility(double igperp, double dg, const double *t,
                                         const double *f, double *out) {
  const double tf = t[0] * f[0] + t[1] * f[1] + t[2] * f[2];
  for (int q = 0; q < 3; ++q) out[q] = igperp * f[q] + dg * t[q] * tf;
}

}  // namespace

template <int G, int S, bool EXACT>
__global__ __launch_bounds__(64) void k_integrate(Dev d, Geo g,
                                                  const int32_t *list,
                                                  int count, int prec) {
  constexpr int NG = 64 / G;
  constexpr int MAXN = G * S;
  const Ctl *ctl = d.ctl;
  if (ctl->abort) return;
  const int lane = threadIdx.x;
  const int grp = lane / G, gl = lane % G;
  const int fidx = blockIdx.x * NG + grp;
  const bool active = fidx < count;
  const int f = active ? list[fidx] : 0;
  const int64_t s0 = active ? d.site_off[f] : 0;
  const int n = active ? (int)(d.site_off[f + 1] - s0) : 0;
  const int64_t b0 = s0 - f;
  const int64_t ns = g.ns, nb = g.nb;
  const bool mid = ctl->midstep != 0;
  const int64_t step = ctl->cur_step;
  double bl = 1.0, kappa = 0.0, gpar = 1.0, gperp = 1.0, sgpar = 0.0, sgperp = 0.0;
  if (active) {
    const double *fp = d.fpar + 8 * f;
    bl = fp[0];
    kappa = fp[1];
    gpar = fp[2];
    gperp = fp[3];
    sgpar = fp[4];
    sgperp = fp[5];
  }

  __shared__ double U_[3][NG][MAXN], F_[3][NG][MAXN], T_[3][NG][MAXN];
  __shared__ double R_[3][NG][MAXN], X_[3][NG][MAXN];
  __shared__ double A_[NG][MAXN], B_[NG][MAXN], C_[NG][MAXN], D_[NG][MAXN];
  __shared__ double CS_[NG][MAXN];
  double *U[3] = {U_[0][grp], U_[1][grp], U_[2][grp]};
  double *FS[3] = {F_[0][grp], F_[1][grp], F_[2][grp]};
  double *TN[3] = {T_[0][grp], T_[1][grp], T_[2][grp]};
  double *RR[3] = {R_[0][grp], R_[1][grp], R_[2][grp]};  // f_rand, then r_new
  double *XX[3] = {X_[0][grp], X_[1][grp], X_[2][grp]};  // scratch
  double *A = A_[grp], *B = B_[grp], *C = C_[grp], *Dd = D_[grp], *CS = CS_[grp];

  // ---- loads (issued up front; per-lane registers) ----------------------
  double rp_reg[S][3];
  double Fb[S][3], Tb[S][3];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i < n) {
#pragma unroll
      for (int q = 0; q < 3; ++q) rp_reg[k][q] = mid ? d.r[q * ns + s0 + i] : d.rp[q * ns + s0 + i];
    }
    if (i < n - 1) {
      const int64_t j = b0 + i;
      double rs[3], ru[3], rl;
      rec_read(d.br, g, j, rs, ru, &rl);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        Fb[k][q] = bond_force(d, nb, q, j);
        Tb[k][q] = bond_torque(d, nb, q, j);
        U[q][i] = ru[q];
      }
    } else if (i == n - 1) {
      double rs[3], ru[3], rl;
      rec_read(d.br, g, b0 + n - 2, rs, ru, &rl);
#pragma unroll
      for (int q = 0; q < 3; ++q) U[q][i] = ru[q];
    }
    if (i < n && !mid && g.stoch) {
#pragma unroll
      for (int q = 0; q < 3; ++q)
        RR[q][i] = g.compact ? (double)reinterpret_cast<const float *>(d.fr)[q * ns + s0 + i]
                             : d.fr[q * ns + s0 + i];
    }
  }
  __syncthreads();

  // ---- ApplyInteractionForces (src/filament.cpp:992-1048) ----------------
  // per bond: sf = F/2 into XX, pt = (u x T)/l into TN (TN is rebuilt below)
  const double linv = 1.0 / bl;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i < n - 1) {
      double uj[3] = {U[0][i], U[1][i], U[2][i]};
      double pt[3];
      cross3(uj, Tb[k], pt);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        TN[q][i] = pt[q] * linv;
        XX[q][i] = 0.5 * Fb[k][q];
      }
    }
  }
  __syncthreads();
  double fsum[S][3];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
#pragma unroll
    for (int q = 0; q < 3; ++q) fsum[k][q] = 0.0;
    if (i < n) {
      if (i >= 1) {
#pragma unroll
        for (int q = 0; q < 3; ++q) fsum[k][q] += XX[q][i - 1];
#pragma unroll
        for (int q = 0; q < 3; ++q) fsum[k][q] += -TN[q][i - 1];
      }
      if (i <= n - 2) {
#pragma unroll
        for (int q = 0; q < 3; ++q) fsum[k][q] += XX[q][i];
#pragma unroll
        for (int q = 0; q < 3; ++q) fsum[k][q] += TN[q][i];
      }
    }
  }
  __syncthreads();
  // ---- CalculateAngles / CalculateTangents ------------------------------
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i >= n) continue;
#pragma unroll
    for (int q = 0; q < 3; ++q) FS[q][i] = fsum[k][q];
    if (i < n - 2) {
      double a[3] = {U[0][i], U[1][i], U[2][i]};
      double b[3] = {U[0][i + 1], U[1][i + 1], U[2][i + 1]};
      CS[i] = dot3(a, b);
    }
    if (i == 0 || i == n - 1) {
      const int bb = (i == 0 ? 0 : n - 2);
#pragma unroll
      for (int q = 0; q < 3; ++q) TN[q][i] = U[q][bb];
    } else {
      double t[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) t[q] = U[q][i - 1] + U[q][i];
      if (EXACT) {
        normalize3(t);
      } else {
        const double inv = 1.0 / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
        for (int q = 0; q < 3; ++q) t[q] *= inv;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) TN[q][i] = t[q];
    }
  }
  __syncthreads();

  // ---- tridiagonal solve helper (EXACT: Thomas on the leader; FAST: PCR) --
  // System rows i < m: lower coefficient in A[i] (x[i-1]), diagonal B[i],
  // upper C[i] (x[i+1]), right-hand side D[i]; EXACT keeps the reference's
  // layout lower[i] = A[i] coupling row i+1.
  auto solve = [&](int m) {
    if (EXACT) {
      if (gl == 0 && m > 0) thomas_reg(m, A, B, C, Dd);
      __syncthreads();
    } else {
      // the round count is the bin's, the same for every group of the
      // block (barriers inside); rounds with s >= m leave a row unchanged.
      // Each round shares 1/B through RB so a row divides once.
      double *RB = XX[2];
      for (int s = 1; s < MAXN; s <<= 1) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
          const int i = gl + G * k;
          if (i < m) RB[i] = 1.0 / B[i];
        }
        __syncthreads();
        double na[S], nbv[S], nc[S], nd[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
          const int i = gl + G * k;
          if (i >= m) continue;
          const double ai = A[i], bi = B[i], ci = C[i], di = Dd[i];
          double A2 = 0.0, C2 = 0.0, B2 = bi, D2 = di;
          if (i - s >= 0) {
            const double al = -ai * RB[i - s];
            A2 = al * A[i - s];
            B2 += al * C[i - s];
            D2 += al * Dd[i - s];
          }
          if (i + s < m) {
            const double ga = -ci * RB[i + s];
            C2 = ga * C[i + s];
            B2 += ga * A[i + s];
            D2 += ga * Dd[i + s];
          }
          na[k] = A2;
          nbv[k] = B2;
          nc[k] = C2;
          nd[k] = D2;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < S; ++k) {
          const int i = gl + G * k;
          if (i >= m) continue;
          A[i] = na[k];
          B[i] = nbv[k];
          C[i] = nc[k];
          Dd[i] = nd[k];
        }
        __syncthreads();
      }
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = gl + G * k;
        if (i < m) Dd[i] = Dd[i] / B[i];
      }
      __syncthreads();
    }
  };
  // coefficients of row i of a symmetric tridiagonal system with
  // off-diagonal o coupling rows i and i+1
  auto put_offdiag = [&](int i, double o, int m) {
    if (EXACT) {
      A[i] = o;  // lower[i]: row i+1, column i
      C[i] = o;  // upper[i]: row i, column i+1
    } else {
      C[i] = o;       // row i, column i+1
      A[i + 1] = o;   // row i+1, column i
      if (i == 0) A[0] = 0.0;
      if (i + 1 == m - 1) C[m - 1] = 0.0;
    }
  };

  if (mid) {
    if (g.stoch) {
      // ConstructUnprojectedRandomForces (src/filament.cpp:518-548)
      const int64_t mstep = (step - 1) >> 1;
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = gl + G * k;
        if (i >= n) continue;
        double xi[3];
        draw_xi(g, (int64_t)d.noise_off[f] + i, mstep, xi);
        const double t0 = TN[0][i], t1 = TN[1][i], t2 = TN[2][i];
        double xt[3];
        if (EXACT) {
          xt[0] = t0 * t0 * xi[0] + t0 * t1 * xi[1] + t0 * t2 * xi[2];
          xt[1] = t1 * t1 * xi[1] + t0 * t1 * xi[0] + t1 * t2 * xi[2];
          xt[2] = t2 * t2 * xi[2] + t0 * t2 * xi[0] + t1 * t2 * xi[1];
        } else {
          const double txi = t0 * xi[0] + t1 * xi[1] + t2 * xi[2];
          xt[0] = t0 * txi;
          xt[1] = t1 * txi;
          xt[2] = t2 * txi;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) RR[q][i] = sgperp * xi[q] + (sgpar - sgperp) * xt[q];
      }
      __syncthreads();
      // GeometricallyProjectRandomForces (src/filament.cpp:550-599)
      const int m = n - 1;
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = gl + G * k;
        if (i >= m) continue;
        double tmp[3], ui[3] = {U[0][i], U[1][i], U[2][i]};
#pragma unroll
        for (int q = 0; q < 3; ++q) tmp[q] = RR[q][i + 1] - RR[q][i];
        Dd[i] = dot3(tmp, ui);
        B[i] = 2;
        if (i < m - 1) put_offdiag(i, -CS[i], m);
      }
      __syncthreads();
      solve(m);
      // the projection of each site reads only its own unprojected force
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int i = gl + G * k;
        if (i >= n) continue;
        double p[3];
        if (i == 0) {
#pragma unroll
          for (int q = 0; q < 3; ++q) p[q] = RR[q][0] + Dd[0] * U[q][0];
        } else if (i == n - 1) {
#pragma unroll
          for (int q = 0; q < 3; ++q) p[q] = RR[q][i] - Dd[n - 2] * U[q][n - 2];
        } else {
#pragma unroll
          for (int q = 0; q < 3; ++q)
            p[q] = RR[q][i] + Dd[i] * U[q][i] - Dd[i - 1] * U[q][i - 1];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (g.compact) {  // mixed precision keeps the random force in FP32
            p[q] = (double)(float)p[q];
            reinterpret_cast<float *>(d.fr)[q * ns + s0 + i] = (float)p[q];
          } else {
            d.fr[q * ns + s0 + i] = p[q];
          }
          RR[q][i] = p[q];
        }
      }
    }
    // Mesh::UpdatePrevPositions (src/mesh.cpp:175-182)
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = gl + G * k;
      if (i < n)
#pragma unroll
        for (int q = 0; q < 3; ++q) d.rp[q * ns + s0 + i] = rp_reg[k][q];
    }
  }
  // AddRandomForces (src/filament.cpp:601-606)
  if (g.stoch) {
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = gl + G * k;
      if (i < n)
#pragma unroll
        for (int q = 0; q < 3; ++q) FS[q][i] += RR[q][i];
    }
  }
  __syncthreads();

  // ---- CalculateBendingForces, 3D (src/filament.cpp:608-776) ------------
  // det_t in A[0..n-1], det_b[j] in C[j-1] for j = 1..n, k_eff in XX[0]
  if (gl == 0 && n > 0) {
    double t2 = 1.0, t1 = 2.0;
    A[0] = 1;
    A[1] = 2;
    double b2 = 1.0, b1 = 2.0;
    C[n - 1] = 1;
    C[n - 2] = 2;
    for (int i = 2; i < n; ++i) {
      const double ca = CS[i - 2];
      const double v = 2 * t1 - (-ca) * (-ca) * t2;
      A[i] = v;
      t2 = t1;
      t1 = v;
      const double cb = CS[n - i - 1];
      const double w = 2 * b1 - (-cb) * (-cb) * b2;
      C[n - i - 1] = w;
      b2 = b1;
      b1 = w;
    }
  }
  __syncthreads();
  double *KE = XX[0];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i < n - 2) {
      if (EXACT) {
        const double ginv = CS[i] * A[i] * C[i + 2] / A[n - 1];
        KE[i] = (kappa + bl * ginv) / (bl * bl);
      } else {
        const double ginv = CS[i] * A[i] * C[i + 2] * (1.0 / A[n - 1]);
        KE[i] = (kappa + bl * ginv) * (linv * linv);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int kk = gl + G * k;
    if (kk >= n) continue;
    double f3[3] = {0.0, 0.0, 0.0};
    if (EXACT) {
      if (kk > 1) {
        double u1[3] = {U[0][kk - 2], U[1][kk - 2], U[2][kk - 2]};
        double u2[3] = {U[0][kk - 1], U[1][kk - 1], U[2][kk - 1]};
        bend_add(f3, KE[kk - 2], u1, u2);
      }
      if (kk > 0 && kk < n - 1) {
        double u1[3] = {U[0][kk - 1], U[1][kk - 1], U[2][kk - 1]};
        double u2[3] = {U[0][kk], U[1][kk], U[2][kk]};
        bend_mid(f3, KE[kk - 1], u1, u2);
      }
      if (kk < n - 2) {
        double u1[3] = {U[0][kk], U[1][kk], U[2][kk]};
        double u2[3] = {U[0][kk + 1], U[1][kk + 1], U[2][kk + 1]};
        bend_sub(f3, KE[kk], u1, u2);
      }
    } else {
      // bond k-1 couples with k-2 and k: f_k = K_{k-2} P(u_{k-1}) u_{k-2}
      //   + K_{k-1} [P(u_{k-1}) u_k - P(u_k) u_{k-1}] - K_k P(u_k) u_{k+1},
      // P(b) a = (I - b b^T) a
      double um2[3], um1[3], u0[3], up1[3], pa[3], pb[3];
      if (kk > 1)
        for (int q = 0; q < 3; ++q) um2[q] = U[q][kk - 2];
      if (kk > 0)
        for (int q = 0; q < 3; ++q) um1[q] = U[q][kk - 1];
      for (int q = 0; q < 3; ++q) u0[q] = U[q][kk];
      if (kk < n - 2)
        for (int q = 0; q < 3; ++q) up1[q] = U[q][kk + 1];
      if (kk > 1) {
        proj_perp(um2, um1, pa);
        const double ke = KE[kk - 2];
        for (int q = 0; q < 3; ++q) f3[q] += ke * pa[q];
      }
      if (kk > 0 && kk < n - 1) {
        proj_perp(u0, um1, pa);
        proj_perp(um1, u0, pb);
        const double ke = KE[kk - 1];
        for (int q = 0; q < 3; ++q) f3[q] += ke * (pa[q] - pb[q]);
      }
      if (kk < n - 2) {
        proj_perp(up1, u0, pa);
        const double ke = KE[kk];
        for (int q = 0; q < 3; ++q) f3[q] -= ke * pa[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) FS[q][kk] += f3[q];
  }
  __syncthreads();

  // ---- CalculateTensions (src/filament.cpp:778-838) ---------------------
  const double igpar = 1.0 / gpar, igperp = 1.0 / gperp;
  {
    const int m = n - 1;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const int i = gl + G * k;
      if (i >= m) continue;
      double t1[3] = {TN[0][i], TN[1][i], TN[2][i]};
      double t2[3] = {TN[0][i + 1], TN[1][i + 1], TN[2][i + 1]};
      double f1[3] = {FS[0][i], FS[1][i], FS[2][i]};
      double f2[3] = {FS[0][i + 1], FS[1][i + 1], FS[2][i + 1]};
      double fd[3];
      if (EXACT) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double ta_ = gam(igpar, igperp, t1, a, 0) * f1[0] + gam(igpar, igperp, t1, a, 1) * f1[1];
          ta_ += gam(igpar, igperp, t1, a, 2) * f1[2];
          double tb_ = gam(igpar, igperp, t2, a, 0) * f2[0] + gam(igpar, igperp, t2, a, 1) * f2[1];
          tb_ += gam(igpar, igperp, t2, a, 2) * f2[2];
          fd[a] = tb_ - ta_;
        }
      } else {
        double m1[3], m2[3];
        mobility(igperp, igpar - igperp, t1, f1, m1);
        mobility(igperp, igpar - igperp, t2, f2, m2);
        for (int a = 0; a < 3; ++a) fd[a] = m2[a] - m1[a];
      }
      double u2[3] = {U[0][i], U[1][i], U[2][i]};
      Dd[i] = dot3(u2, fd);
      const double a1 = dot3(t1, u2), a2 = dot3(t2, u2);
      B[i] = 2.0 / gperp + (1.0 / gpar - 1.0 / gperp) * (a1 * a1 + a2 * a2);
      if (i > 0) {
        double u1[3] = {U[0][i - 1], U[1][i - 1], U[2][i - 1]};
        const double hu = -1.0 / gperp * dot3(u2, u1) -
                          (1.0 / gpar - 1.0 / gperp) * (dot3(t1, u1) * dot3(t1, u2));
        put_offdiag(i - 1, hu, m);
      }
    }
    __syncthreads();
    solve(m);
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i < n - 1) d.tension[b0 + i] = Dd[i];
  }

  // ---- UpdateSitePositions (src/filament.cpp:840-927) -------------------
  const double delta = mid ? 0.5 * g.delta : g.delta;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i >= n) continue;
    double f3[3] = {FS[0][i], FS[1][i], FS[2][i]};
    if (i == 0) {
#pragma unroll
      for (int q = 0; q < 3; ++q) f3[q] += Dd[0] * U[q][0];
    } else if (i == n - 1) {
#pragma unroll
      for (int q = 0; q < 3; ++q) f3[q] += -Dd[n - 2] * U[q][n - 2];
    } else {
#pragma unroll
      for (int q = 0; q < 3; ++q) f3[q] += Dd[i] * U[q][i] - Dd[i - 1] * U[q][i - 1];
    }
    double t3[3] = {TN[0][i], TN[1][i], TN[2][i]};
    if (EXACT) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double ft = gam(igpar, igperp, t3, a, 0) * f3[0] + gam(igpar, igperp, t3, a, 1) * f3[1];
        ft += gam(igpar, igperp, t3, a, 2) * f3[2];
        RR[a][i] = rp_reg[k][a] + delta * ft;
      }
    } else {
      double mv[3];
      mobility(igperp, igpar - igperp, t3, f3, mv);
#pragma unroll
      for (int a = 0; a < 3; ++a) RR[a][i] = rp_reg[k][a] + delta * mv[a];
    }
  }
  __syncthreads();
  // new orientations in XX, CheckBondLengths (src/filament.cpp:934-951)
  int bad = 0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i >= n - 1) continue;
    double rd[3], mm = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      rd[q] = RR[q][i + 1] - RR[q][i];
      mm += rd[q] * rd[q];
    }
    mm = sqrt(mm);
    if (EXACT) {
#pragma unroll
      for (int q = 0; q < 3; ++q) XX[q][i] = rd[q] / mm;
    } else {
      const double inv = 1.0 / mm;
#pragma unroll
      for (int q = 0; q < 3; ++q) XX[q][i] = rd[q] * inv;
    }
    const double diff = bl - mm;
    const double err = (diff < 0 ? -diff : diff) / bl;
    if (err > 1e-3) bad = 1;
  }
  // per-group OR of the renormalisation flag
  const unsigned long long bal = __ballot(bad);
  const unsigned long long gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << (grp * G));
  const bool renorm = (bal & gmask) != 0;
  __syncthreads();
  if (gl == 0 && n > 0) {
    int32_t sw = d.norm[2 * f], nn = d.norm[2 * f + 1];
    if (renorm) {
      sw = !sw;
      nn = 0;
      if (sw) {
        double x = RR[0][0], y = RR[1][0], z = RR[2][0];
        for (int i = 1; i < n; ++i) {
          x = x + bl * XX[0][i - 1];
          y = y + bl * XX[1][i - 1];
          z = z + bl * XX[2][i - 1];
          RR[0][i] = x;
          RR[1][i] = y;
          RR[2][i] = z;
        }
      } else {
        double x = RR[0][n - 1], y = RR[1][n - 1], z = RR[2][n - 1];
        for (int i = n - 1; i > 0; --i) {
          x = x - bl * XX[0][i - 1];
          y = y - bl * XX[1][i - 1];
          z = z - bl * XX[2][i - 1];
          RR[0][i - 1] = x;
          RR[1][i - 1] = y;
          RR[2][i - 1] = z;
        }
      }
    }
    nn++;
    d.norm[2 * f] = sw;
    d.norm[2 * f + 1] = nn;
  }
  __syncthreads();

  // ---- write back sites, Mesh::UpdateBondPositions, rebuild bookkeeping --
  double lmax = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const int i = gl + G * k;
    if (i >= n) continue;
    double rn[3] = {RR[0][i], RR[1][i], RR[2][i]};
#pragma unroll
    for (int q = 0; q < 3; ++q) d.r[q * ns + s0 + i] = rn[q];
    if (i < n - 1) {
      double r2[3] = {RR[0][i + 1], RR[1][i + 1], RR[2][i + 1]};
      bond_reinit(d, g, b0 + i, rn, r2, prec);
    }
    if (!mid) {
      // Mesh::UpdateDrTot for the next step's rebuild test, two
      // Object::UpdateDrTot calls per site (src/mesh.cpp:533-544,
      // src/object.cpp:274-287)
      double a = d.acc[s0 + i];
      d.acc_prev[s0 + i] = a;
      for (int rep = 0; rep < 2; ++rep)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const double dd = rn[q] - d.dr0[q * ns + s0 + i];
          a += dd * dd;
        }
      d.acc[s0 + i] = a;
      lmax = a > lmax ? a : lmax;
    }
  }
  if (!mid) {
#pragma unroll
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

static int integrate_bin_lanes(int maxn) {
  // lanes per filament: the bin size up to 64, then 64 with 2 or 4 sites
  // per lane
  return maxn <= 64 ? maxn : 64;
}

void launch_integrate(const Dev &d, const Geo &g, const int32_t *bin_list,
                      int bin_count, int bin_maxn, int precision, const Tune &t, bool mid,
                      hipStream_t st) {
  if (bin_count <= 0) return;
  // fast path: register-resident variants
  if (launch_integrate_long(d, g, bin_list, bin_count, bin_maxn, precision, t.int_waves,
                            t.int_layout, mid, st))
    return;
  if (launch_integrate_reg(d, g, bin_list, bin_count, bin_maxn, precision, t.int_waves, st))
    return;
  const bool exact = precision == SC_PREC_FP64;
  const int G = integrate_bin_lanes(bin_maxn);
  const int per_wave = 64 / G;
  dim3 grid((bin_count + per_wave - 1) / per_wave), block(64);
#define SC_LAUNCH(GG, SS)                                                          \
  do {                                                                             \
    if (exact)                                                                     \
      hipLaunchKernelGGL((k_integrate<GG, SS, true>), grid, block, 0, st, d, g,    \
                         bin_list, bin_count, precision);                          \
    else                                                                           \
      hipLaunchKernelGGL((k_integrate<GG, SS, false>), grid, block, 0, st, d, g,   \
                         bin_list, bin_count, precision);                          \
  } while (0)
  switch (bin_maxn) {
    case 8: SC_LAUNCH(8, 1); break;
    case 16: SC_LAUNCH(16, 1); break;
    case 32: SC_LAUNCH(32, 1); break;
    case 64: SC_LAUNCH(64, 1); break;
    case 128: SC_LAUNCH(64, 2); break;
    default: SC_LAUNCH(64, 4); break;
  }
#undef SC_LAUNCH
}

}  // namespace sc
