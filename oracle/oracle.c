/*
 * oracle.c -- CPU restatement of simcore's filament BD hot path (test
 * infrastructure only; see oracle.h).  Build with -ffp-contract=off so that
 * every multiply and add rounds separately, as the reference's do when built
 * for x86-64 without FMA.
 *
 * Parity pins, per function (tests/test_oracle_*.py):
 *   philox4x32          Random123 known-answer vectors
 *   cell_key            hand-picked faces / NINT half-way cases
 *   cell_neighbors      brute force: every adjacent (periodic) cell pair
 *                       covered exactly once, 3^3..7^3 grids
 *   rebuild/emission    brute force O(N^2) candidate set on random systems
 *   min_distance        dense parameter sampling + golden-section refine,
 *                       closed forms (crossing, parallel, collinear)
 *   pair_force          analytic WCA: F(sigma)=24 eps/sigma, F(rc)=0, clamp,
 *                       finite-difference -dU/dr; max-force fallback cases
 *   tridiag             dense LU (numpy) on random SPD tridiagonal systems
 *   bending_forces      straight filament => 0, det_t[i]=i+1, g_inv from
 *                       dense determinant ratios, kappa part = -grad of
 *                       (kappa/l) sum (1-u_i.u_j) by finite differences
 *   random projection   u_i.(f_{i+1}-f_i) = 0 after projection
 *   integrator step     first-order constraint residual, bond lengths kept,
 *                       free-filament diffusion statistics
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/simcore/macros.hpp:6,7,14,15 */
#define O_NINT(x) ((x) < 0.0 ? (int)((x)-0.5) : (int)((x) + 0.5))
#define O_ABS(x) ((x) < 0 ? -(x) : (x))
#define O_SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))
#define O_SIGNOF(x) ((x) >= 0.0 ? 1 : -1)
#define O_SMALL 1.0e-12 /* src/minimum_distance.cpp:3 */

struct orc_sys {
  orc_params p;
  double c12, c6, rcut2;
  int n1d;
  double cell_len, dr_update;
  int nf;
  int64_t ns, nb;
  int32_t *site_off, *bond_fil, *bond_idx; /* bond -> filament, index in fil */
  int max_sites;
  /* per filament */
  double *bl, *kappa, *gpar, *gperp, *diam, *sig_par, *sig_perp, *dr_tot;
  unsigned char *norm_switch;
  int *n_normalize;
  /* sites (SoA) */
  double *r[3], *rp[3], *fr[3], *dr0[3], *acc;
  /* bonds (SoA) */
  double *pos[3], *sp[3], *u[3], *len, *bf[3], *bt[3], *tension;
  int32_t *key;
  /* cells */
  int64_t n_cells;
  int32_t *cell_start, *cell_obj;
  /* candidates */
  int64_t nc, cap;
  int32_t *c1, *c2;
  unsigned char *no_ix;
  double *cf[3], *ct1[3], *ct2[3];
  /* Interaction::stress / pote of each pair (sticky like the force), the
   * InteractionManager's stress_ sum and each bond's p_energy_ */
  double *cst[9], *cpote;
  double stress_acc[9];
  double *bpote;
  /* state */
  int64_t rebuilds, i_step, n_midsteps;
  /* caller-supplied noise: xi of midsteps [xi_m0, xi_m0 + xi_nm), 3 per
   * site, site-major (NULL: the counter-based stream) */
  double *xi_buf;
  int64_t xi_m0, xi_nm;
  int i_update;
  int midstep;        /* FilamentSpecies::midstep_ (next update's flag) */
  int last_midstep;   /* Mesh::midstep_ (flag of the last update)       */
  int err;
  /* per-thread scratch */
  int n_thr;
  double **scratch;
  /* crosslinks */
  int64_t nx, xl_events;
  int32_t *xb;     /* 2 per crosslink */
  double *xlam;    /* 2 per crosslink */
  double *xf;      /* 9 per crosslink: F, T0, T1 */
  orc_xlink_params xp;
  double xpoff;
  int xreach;
};

/* ---------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the noise injection.
 * ------------------------------------------------------------------------- */
static inline void ph_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    ph_round(c, k);
  }
  memcpy(out, c, sizeof(c));
}

/* xi_d = (x_d + 1/2) 2^-32 - 1/2, exact in FP64; replaces
 * rng_.RandomUniform() - 0.5 (src/filament.cpp:529-530). */
void orc_noise(uint64_t seed, int64_t site, int64_t m, double xi[3]) {
  uint32_t ctr[4] = {(uint32_t)site, (uint32_t)((uint64_t)site >> 32),
                     (uint32_t)m, (uint32_t)((uint64_t)m >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32(ctr, key, o);
  for (int d = 0; d < 3; ++d)
    xi[d] = ((double)o[d] + 0.5) * 2.3283064365386962890625e-10 - 0.5;
}

/* ---------------------------------------------------------------------------
 * Small vector helpers with the reference's summation order
 * (src/linear_algebra.cpp:3-35).
 * ------------------------------------------------------------------------- */
static inline double dot3(const double *a, const double *b) {
  double m = 0.0;
  m += a[0] * b[0];
  m += a[1] * b[1];
  m += a[2] * b[2];
  return m;
}
static inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Filament::ApplyInteractionForces (src/filament.cpp:992-1048), driving
 * zero: bond i (orientation u, force F, torque T) adds F/2 + (u x T)/l to
 * site i and F/2 - (u x T)/l to site i+1.  u, F, T per bond as three
 * component arrays; fs accumulated per site, AoS (3 per site). */
static void bond_site_forces(int n, double bl, const double *const u3[3],
                             const double *const f3[3], const double *const t3[3],
                             double *fs) {
  double linv = 1.0 / bl;
  for (int i = 0; i < n - 1; ++i) {
    double F[3] = {f3[0][i], f3[1][i], f3[2][i]};
    double T[3] = {t3[0][i], t3[1][i], t3[2][i]};
    double u[3] = {u3[0][i], u3[1][i], u3[2][i]};
    double pt[3], sf[3];
    cross3(u, T, pt);
    for (int d = 0; d < 3; ++d) {
      pt[d] *= linv;
      sf[d] = 0.5 * F[d];
    }
    for (int d = 0; d < 3; ++d) fs[3 * i + d] += sf[d];
    for (int d = 0; d < 3; ++d) fs[3 * i + d] += pt[d];
    for (int d = 0; d < 3; ++d) pt[d] *= -1;
    for (int d = 0; d < 3; ++d) fs[3 * (i + 1) + d] += sf[d];
    for (int d = 0; d < 3; ++d) fs[3 * (i + 1) + d] += pt[d];
  }
}

void orc_bond_site_forces(int n_sites, double bond_length, const double *u, const double *f,
                          const double *t, double *f_site) {
  const int nb = n_sites - 1;
  const double *U[3] = {u, u + nb, u + 2 * nb};
  const double *F[3] = {f, f + nb, f + 2 * nb};
  const double *T[3] = {t, t + nb, t + 2 * nb};
  double *fs = (double *)calloc((size_t)3 * n_sites, sizeof(double));
  if (!fs) return;
  bond_site_forces(n_sites, bond_length, U, F, T, fs);
  for (int i = 0; i < n_sites; ++i)
    for (int d = 0; d < 3; ++d) f_site[d * n_sites + i] = fs[3 * i + d];
  free(fs);
}
static inline void normalize3(double *a) {
  double m = 0.0;
  m += a[0] * a[0];
  m += a[1] * a[1];
  m += a[2] * a[2];
  m = sqrt(m);
  a[0] /= m;
  a[1] /= m;
  a[2] /= m;
}

/* tridiagonal_solver, src/linear_algebra.cpp:164-197 */
void orc_tridiag(int n, const double *a, const double *b, double *c,
                 double *d) {
  n--;
  c[0] /= b[0];
  d[0] /= b[0];
  for (int i = 1; i < n; ++i) {
    c[i] /= b[i] - a[i - 1] * c[i - 1];
    d[i] = (d[i] - a[i - 1] * d[i - 1]) / (b[i] - a[i - 1] * c[i - 1]);
  }
  d[n] = (d[n] - a[n - 1] * d[n - 1]) / (b[n] - a[n - 1] * c[n - 1]);
  for (int i = n; i-- > 0;) d[i] -= c[i] * d[i + 1];
}

/* ---------------------------------------------------------------------------
 * Cell list: CellList::FindCellCoords (src/cell_list.cpp:139-157) and the
 * non-redundant half-shell of AssignCellNeighbors(false) (:206-262), for a
 * fully periodic 3D box.  Keys follow MakePairs' loop order cell_[x][y][z]
 * (:127-137): key = (x*n + y)*n + z.
 * ------------------------------------------------------------------------- */
int32_t orc_cell_key(int n1d, const double *spos) {
  int c[3];
  for (int d = 0; d < 3; ++d) {
    double x = spos[d] + 0.5;
    int xc = (int)floor(n1d * x);
    if (xc == n1d) xc -= 1;
    /* s = -0.5 - 2^-54 (NINT of a value just below 0.5 rounding up) would
     * give cell -1, an out-of-range index in the reference; clamp. */
    if (xc < 0) xc = 0;
    if (xc >= n1d) xc = n1d - 1;
    c[d] = xc;
  }
  return ((int32_t)c[0] * n1d + c[1]) * n1d + c[2];
}

int orc_cell_neighbors(int n1d, int x, int y, int z, int32_t *out) {
  int cnt = 0;
  for (int zp = z; zp < z + 2; ++zp) {
    int nz = zp;
    if (nz < 0 || nz == n1d) nz = (nz < 0 ? n1d - 1 : 0);
    int y_begin = (zp > z ? y - 1 : y);
    for (int yp = y_begin; yp < y + 2; ++yp) {
      int ny = yp;
      if (ny < 0 || ny == n1d) ny = (ny < 0 ? n1d - 1 : 0);
      int x_begin = (yp > y || zp > z ? x - 1 : x + 1);
      for (int xp = x_begin; xp < x + 2; ++xp) {
        int nx = xp;
        if (nx < 0 || nx == n1d) nx = (nx < 0 ? n1d - 1 : 0);
        if (nx == x && ny == y && nz == z) continue;
        out[cnt++] = ((int32_t)nx * n1d + ny) * n1d + nz;
      }
    }
  }
  return cnt;
}

/* ---------------------------------------------------------------------------
 * MinimumDistance::Sphero for a fully periodic box
 * (src/minimum_distance.cpp:202-344).
 * ------------------------------------------------------------------------- */
static void sphero(const double *h, const double *s_1, const double *u_1,
                   double length_1, const double *s_2, const double *u_2,
                   double length_2, double *r_min, double *r_min_mag2,
                   double *contact_1, double *contact_2) {
  double dr[3], ds[3];
  double half_length_1 = 0.5 * length_1;
  double half_length_2 = 0.5 * length_2;
  for (int i = 0; i < 3; ++i) {
    ds[i] = s_2[i] - s_1[i];
    ds[i] -= O_NINT(ds[i]);
  }
  for (int i = 0; i < 3; ++i) {
    dr[i] = 0.0;
    for (int j = 0; j < 3; ++j) dr[i] += h[3 * i + j] * ds[j];
  }
  double dr_dot_u_1 = 0.0, dr_dot_u_2 = 0.0, u_1_dot_u_2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    dr_dot_u_1 += dr[i] * u_1[i];
    dr_dot_u_2 += dr[i] * u_2[i];
    u_1_dot_u_2 += u_1[i] * u_2[i];
  }
  double lambda, mu;
  double denom = 1.0 - u_1_dot_u_2 * u_1_dot_u_2;
  if (denom < O_SMALL) {
    lambda = dr_dot_u_1 / 2.0;
    mu = -dr_dot_u_2 / 2.0;
  } else {
    lambda = (dr_dot_u_1 - u_1_dot_u_2 * dr_dot_u_2) / denom;
    mu = (-dr_dot_u_2 + u_1_dot_u_2 * dr_dot_u_1) / denom;
  }
  double lambda_mag = O_ABS(lambda);
  double mu_mag = O_ABS(mu);
  if (lambda_mag > half_length_1 && mu_mag > half_length_2) {
    double lambda_a = O_SIGN(half_length_1, lambda);
    double mu_a = -dr_dot_u_2 + lambda_a * u_1_dot_u_2;
    mu_mag = O_ABS(mu_a);
    if (mu_mag > half_length_2) mu_a = O_SIGN(half_length_2, mu_a);
    double ra[3], rb[3], m2a = 0.0, m2b = 0.0;
    for (int i = 0; i < 3; ++i) {
      ra[i] = dr[i] - lambda_a * u_1[i] + mu_a * u_2[i];
      m2a += ra[i] * ra[i];
    }
    double mu_b = O_SIGN(half_length_2, mu);
    double lambda_b = dr_dot_u_1 + mu_b * u_1_dot_u_2;
    lambda_mag = O_ABS(lambda_b);
    if (lambda_mag > half_length_1) lambda_b = O_SIGN(half_length_1, lambda_b);
    for (int i = 0; i < 3; ++i) {
      rb[i] = dr[i] - lambda_b * u_1[i] + mu_b * u_2[i];
      m2b += rb[i] * rb[i];
    }
    if (m2a < m2b) {
      lambda = lambda_a;
      mu = mu_a;
      *r_min_mag2 = m2a;
      for (int i = 0; i < 3; ++i) r_min[i] = ra[i];
    } else {
      lambda = lambda_b;
      mu = mu_b;
      *r_min_mag2 = m2b;
      for (int i = 0; i < 3; ++i) r_min[i] = rb[i];
    }
  } else if (lambda_mag > half_length_1) {
    lambda = O_SIGN(half_length_1, lambda);
    mu = -dr_dot_u_2 + lambda * u_1_dot_u_2;
    mu_mag = O_ABS(mu);
    if (mu_mag > half_length_2) mu = O_SIGN(half_length_2, mu);
    *r_min_mag2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      r_min[i] = dr[i] - lambda * u_1[i] + mu * u_2[i];
      *r_min_mag2 += r_min[i] * r_min[i];
    }
  } else if (mu_mag > half_length_2) {
    mu = O_SIGN(half_length_2, mu);
    lambda = dr_dot_u_1 + mu * u_1_dot_u_2;
    lambda_mag = O_ABS(lambda);
    if (lambda_mag > half_length_1) lambda = O_SIGN(half_length_1, lambda);
    *r_min_mag2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      r_min[i] = dr[i] - lambda * u_1[i] + mu * u_2[i];
      *r_min_mag2 += r_min[i] * r_min[i];
    }
  } else {
    *r_min_mag2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      r_min[i] = dr[i] - lambda * u_1[i] + mu * u_2[i];
      *r_min_mag2 += r_min[i] * r_min[i];
    }
  }
  for (int i = 0; i < 3; ++i) {
    contact_1[i] = lambda * u_1[i];
    contact_2[i] = mu * u_2[i];
  }
}

void orc_min_distance(const double *h, const double *s1, const double *u1,
                      double l1, const double *s2, const double *u2, double l2,
                      double *out) {
  sphero(h, s1, u1, l1, s2, u2, l2, out, out + 3, out + 4, out + 7);
}

/* PotentialManager::CalcPotential -> WCAPotential / MaxForcePotential
 * (include/simcore/potential_manager.hpp:92-103,
 *  include/simcore/wca_potential.hpp:14-32,
 *  include/simcore/max_force_potential.hpp:11-40).  `dr` may be rewritten
 * by the max-force branch, exactly as ix.dr is. Returns 1 if force written. */
static int pair_force(double c12, double c6, double fcut, double *dr,
                      double dr_mag2, const double *c1, const double *c2,
                      double *force) {
  if (dr_mag2 < 1e-12) {
    if (c1[0] || c1[1] || c1[2]) {
      for (int i = 0; i < 3; ++i) dr[i] = c1[i] - c2[i];
    }
    double rmag = 0.0;
    for (int i = 0; i < 3; ++i) rmag += dr[i] * dr[i];
    rmag = sqrt(rmag);
    if (rmag < 1e-12) return 0;
    double rinv = 1.0 / (rmag);
    for (int i = 0; i < 3; ++i) force[i] = fcut * dr[i] * rinv;
    return 1;
  }
  double rmag = sqrt(dr_mag2);
  double rinv = 1.0 / (rmag);
  double rinv2 = rinv * rinv;
  double r6 = rinv2 * rinv2 * rinv2;
  double ffac = -(12.0 * c12 * r6 - 6.0 * c6) * r6 * rinv;
  if (O_ABS(ffac) > fcut) ffac = O_SIGNOF(ffac) * fcut;
  for (int i = 0; i < 3; ++i) force[i] = ffac * dr[i] * rinv;
  return 1;
}

int orc_pair_force(double c12, double c6, double rcut2, double f_cutoff,
                   const double *dr_in, double dr_mag2, const double *c1,
                   const double *c2, double *force) {
  double dr[3] = {dr_in[0], dr_in[1], dr_in[2]};
  if (dr_mag2 > rcut2) return 0;
  return pair_force(c12, c6, f_cutoff, dr, dr_mag2, c1, c2, force);
}

/* ---------------------------------------------------------------------------
 * System setup
 * ------------------------------------------------------------------------- */
static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  return p;
}

/* Bond::ReInit + Object::UpdatePeriodic (src/bond.cpp:22-37,
 * src/object.cpp:202-210, src/linear_algebra.cpp:100-117). */
static void bond_reinit(orc_sys *s, int64_t b, int64_t site) {
  double o[3], l = 0.0;
  for (int i = 0; i < 3; ++i) {
    o[i] = s->r[i][site + 1] - s->r[i][site];
    l += o[i] * o[i];
  }
  l = sqrt(l);
  double pos[3];
  for (int i = 0; i < 3; ++i) {
    pos[i] = s->r[i][site] + 0.5 * o[i];
    o[i] /= l;
  }
  for (int i = 0; i < 3; ++i) {
    double si = 0.0;
    for (int j = 0; j < 3; ++j) si += s->p.h_inv[3 * i + j] * pos[j];
    si -= O_NINT(si);
    s->sp[i][b] = si;
    s->pos[i][b] = pos[i];
    s->u[i][b] = o[i];
  }
  s->len[b] = l;
}

static void zero_dr_tot(orc_sys *s) {
  /* Mesh::ZeroDrTot / Object::ZeroDrTot (src/mesh.cpp:546-551,
   * src/object.cpp:278-281) */
  for (int i = 0; i < 3; ++i)
    memcpy(s->dr0[i], s->r[i], sizeof(double) * (size_t)s->ns);
  memset(s->acc, 0, sizeof(double) * (size_t)s->ns);
  memset(s->dr_tot, 0, sizeof(double) * (size_t)s->nf);
}

static int ensure_cap(orc_sys *s, int64_t need) {
  if (need <= s->cap) return 0;
  int64_t nc = s->cap ? s->cap : 1024;
  while (nc < need) nc *= 2;
  s->c1 = realloc(s->c1, sizeof(int32_t) * nc);
  s->c2 = realloc(s->c2, sizeof(int32_t) * nc);
  s->no_ix = realloc(s->no_ix, (size_t)nc);
  for (int i = 0; i < 3; ++i) {
    s->cf[i] = realloc(s->cf[i], sizeof(double) * nc);
    s->ct1[i] = realloc(s->ct1[i], sizeof(double) * nc);
    s->ct2[i] = realloc(s->ct2[i], sizeof(double) * nc);
  }
  for (int i = 0; i < 9; ++i) s->cst[i] = realloc(s->cst[i], sizeof(double) * nc);
  s->cpote = realloc(s->cpote, sizeof(double) * nc);
  s->cap = nc;
  return 0;
}

/* InteractionManager::UpdateInteractions -> UpdatePairInteractions ->
 * CellList::RenewObjectsCells + MakePairs, then ZeroDrTot
 * (src/interaction_manager.cpp:211-231, 316-321; src/cell_list.cpp:127-199;
 * src/cell.cpp:14-70).  Interactors are the bonds in species -> member ->
 * bond order (src/interaction_manager.cpp:122-138, src/mesh.cpp:383-397),
 * so pushing bonds in index order into each cell reproduces Cell::AddObj's
 * order; MakePairs then walks cells in key order. */
static void rebuild(orc_sys *s) {
  int64_t ncell = s->n_cells;
  int n = s->n1d;
  memset(s->cell_start, 0, sizeof(int32_t) * (size_t)(ncell + 1));
  for (int64_t b = 0; b < s->nb; ++b) {
    double sp[3] = {s->sp[0][b], s->sp[1][b], s->sp[2][b]};
    int32_t k = orc_cell_key(n, sp);
    s->key[b] = k;
    s->cell_start[k + 1]++;
  }
  for (int64_t c = 0; c < ncell; ++c) s->cell_start[c + 1] += s->cell_start[c];
  {
    int32_t *fill = xcalloc((size_t)ncell, sizeof(int32_t));
    for (int64_t b = 0; b < s->nb; ++b) {
      int32_t k = s->key[b];
      s->cell_obj[s->cell_start[k] + fill[k]++] = (int32_t)b;
    }
    free(fill);
  }
  s->nc = 0;
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      for (int z = 0; z < n; ++z) {
        int32_t c = ((int32_t)x * n + y) * n + z;
        int32_t b0 = s->cell_start[c], m = s->cell_start[c + 1] - b0;
        if (m == 0) continue;
        int32_t nbr[13];
        int nn = orc_cell_neighbors(n, x, y, z, nbr);
        int64_t need = (int64_t)m * (m - 1) / 2;
        for (int k = 0; k < nn; ++k)
          need += (int64_t)m * (s->cell_start[nbr[k] + 1] - s->cell_start[nbr[k]]);
        ensure_cap(s, s->nc + need);
        for (int i = 0; i < m - 1; ++i)
          for (int j = i + 1; j < m; ++j) {
            s->c1[s->nc] = s->cell_obj[b0 + i];
            s->c2[s->nc] = s->cell_obj[b0 + j];
            s->nc++;
          }
        for (int k = 0; k < nn; ++k) {
          int32_t o0 = s->cell_start[nbr[k]];
          int32_t om = s->cell_start[nbr[k] + 1] - o0;
          for (int i = 0; i < m; ++i)
            for (int j = 0; j < om; ++j) {
              s->c1[s->nc] = s->cell_obj[b0 + i];
              s->c2[s->nc] = s->cell_obj[o0 + j];
              s->nc++;
            }
        }
      }
  /* new Interaction records: force/torques zero, flags clear */
  for (int i = 0; i < 3; ++i) {
    memset(s->cf[i], 0, sizeof(double) * (size_t)s->nc);
    memset(s->ct1[i], 0, sizeof(double) * (size_t)s->nc);
    memset(s->ct2[i], 0, sizeof(double) * (size_t)s->nc);
  }
  for (int i = 0; i < 9; ++i) memset(s->cst[i], 0, sizeof(double) * (size_t)s->nc);
  memset(s->cpote, 0, sizeof(double) * (size_t)s->nc);
  memset(s->no_ix, 0, (size_t)s->nc);
  s->i_update = 0;
  s->rebuilds++;
  zero_dr_tot(s);
}

orc_sys *orc_create(const orc_params *p, int nf, const int32_t *site_off,
                    const double *rx, const double *ry, const double *rz,
                    const double *bond_length, const double *kappa,
                    const double *gamma_par, const double *gamma_perp,
                    const double *diameter) {
  if (!p || nf <= 0 || !site_off) return NULL;
  orc_sys *s = xcalloc(1, sizeof(orc_sys));
  s->p = *p;
  /* WCAPotential::Init (include/simcore/wca_potential.hpp:34-48) */
  double rcut = pow(2.0, 1.0 / 6.0) * p->wca_sig;
  s->rcut2 = rcut * rcut;
  s->c12 = 4.0 * p->wca_eps * pow(p->wca_sig, 12.0);
  s->c6 = 4.0 * p->wca_eps * pow(p->wca_sig, 6.0);
  /* InitInteractions: SetMinCellLength(sqrt(rcut2)) then CellList::Init
   * (src/interaction_manager.cpp:46-50, src/cell_list.cpp:9-64) */
  double min_len = p->min_cell_length;
  if (sqrt(s->rcut2) > min_len) min_len = sqrt(s->rcut2);
  s->n1d = (int)floor(2 * p->system_radius / min_len);
  if (s->n1d <= 3) s->n1d = 3;
  s->cell_len = (double)2 * p->system_radius / s->n1d;
  s->dr_update = 0.25 * s->cell_len * s->cell_len;
  s->n_cells = (int64_t)s->n1d * s->n1d * s->n1d;
  s->nf = nf;
  s->site_off = xcalloc((size_t)nf + 1, sizeof(int32_t));
  memcpy(s->site_off, site_off, sizeof(int32_t) * ((size_t)nf + 1));
  s->ns = site_off[nf];
  s->nb = s->ns - nf;
  s->max_sites = 0;
  for (int f = 0; f < nf; ++f) {
    int nsf = site_off[f + 1] - site_off[f];
    if (nsf < 3) {
      free(s->site_off);
      free(s);
      return NULL;
    }
    if (nsf > s->max_sites) s->max_sites = nsf;
  }
  int64_t ns = s->ns, nb = s->nb;
  s->bond_fil = xcalloc((size_t)nb, sizeof(int32_t));
  s->bpote = xcalloc((size_t)nb, sizeof(double));
  s->bond_idx = xcalloc((size_t)nb, sizeof(int32_t));
  double **fp[] = {&s->bl, &s->kappa, &s->gpar, &s->gperp, &s->diam,
                   &s->sig_par, &s->sig_perp, &s->dr_tot};
  for (size_t i = 0; i < sizeof(fp) / sizeof(fp[0]); ++i)
    *fp[i] = xcalloc((size_t)nf, sizeof(double));
  s->norm_switch = xcalloc((size_t)nf, 1);
  s->n_normalize = xcalloc((size_t)nf, sizeof(int));
  for (int i = 0; i < 3; ++i) {
    s->r[i] = xcalloc((size_t)ns, sizeof(double));
    s->rp[i] = xcalloc((size_t)ns, sizeof(double));
    s->fr[i] = xcalloc((size_t)ns, sizeof(double));
    s->dr0[i] = xcalloc((size_t)ns, sizeof(double));
    s->pos[i] = xcalloc((size_t)nb, sizeof(double));
    s->sp[i] = xcalloc((size_t)nb, sizeof(double));
    s->u[i] = xcalloc((size_t)nb, sizeof(double));
    s->bf[i] = xcalloc((size_t)nb, sizeof(double));
    s->bt[i] = xcalloc((size_t)nb, sizeof(double));
  }
  s->acc = xcalloc((size_t)ns, sizeof(double));
  s->len = xcalloc((size_t)nb, sizeof(double));
  s->tension = xcalloc((size_t)nb, sizeof(double));
  s->key = xcalloc((size_t)nb, sizeof(int32_t));
  s->cell_start = xcalloc((size_t)s->n_cells + 1, sizeof(int32_t));
  s->cell_obj = xcalloc((size_t)nb, sizeof(int32_t));
  memcpy(s->r[0], rx, sizeof(double) * (size_t)ns);
  memcpy(s->r[1], ry, sizeof(double) * (size_t)ns);
  memcpy(s->r[2], rz, sizeof(double) * (size_t)ns);
  for (int f = 0; f < nf; ++f) {
    s->bl[f] = bond_length[f];
    s->kappa[f] = kappa[f];
    s->gpar[f] = gamma_par[f];
    s->gperp[f] = gamma_perp[f];
    s->diam[f] = diameter ? diameter[f] : 1.0;
    /* Filament::SetDiffusion (src/filament.cpp:346-347) */
    s->sig_perp[f] = sqrt(24.0 * gamma_perp[f] / p->delta);
    s->sig_par[f] = sqrt(24.0 * gamma_par[f] / p->delta);
    s->norm_switch[f] = 1; /* include/simcore/filament.hpp: normalize_switch_ */
    int b0 = site_off[f] - f;
    for (int i = 0; i < site_off[f + 1] - site_off[f] - 1; ++i) {
      s->bond_fil[b0 + i] = f;
      s->bond_idx[b0 + i] = i;
      bond_reinit(s, b0 + i, site_off[f] + i);
    }
  }
  for (int i = 0; i < 3; ++i)
    memcpy(s->rp[i], s->r[i], sizeof(double) * (size_t)ns);
#ifdef _OPENMP
  s->n_thr = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
#else
  s->n_thr = 1;
#endif
  s->scratch = xcalloc((size_t)s->n_thr, sizeof(double *));
  for (int t = 0; t < s->n_thr; ++t)
    s->scratch[t] = xcalloc((size_t)(s->max_sites + 2) * 40, sizeof(double));
  s->midstep = 1;
  s->last_midstep = 1;
  rebuild(s);
  return s;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->site_off);
  free(s->bond_fil);
  free(s->bond_idx);
  free(s->bl);
  free(s->kappa);
  free(s->gpar);
  free(s->gperp);
  free(s->diam);
  free(s->sig_par);
  free(s->sig_perp);
  free(s->dr_tot);
  free(s->xi_buf);
  free(s->norm_switch);
  free(s->n_normalize);
  for (int i = 0; i < 3; ++i) {
    free(s->r[i]);
    free(s->rp[i]);
    free(s->fr[i]);
    free(s->dr0[i]);
    free(s->pos[i]);
    free(s->sp[i]);
    free(s->u[i]);
    free(s->bf[i]);
    free(s->bt[i]);
    free(s->cf[i]);
    free(s->ct1[i]);
    free(s->ct2[i]);
  }
  free(s->acc);
  free(s->len);
  free(s->tension);
  free(s->key);
  free(s->cell_start);
  free(s->cell_obj);
  free(s->c1);
  free(s->c2);
  free(s->no_ix);
  for (int k = 0; k < 9; ++k) free(s->cst[k]);
  free(s->cpote);
  free(s->bpote);
  free(s->xb);
  free(s->xlam);
  free(s->xf);
  for (int t = 0; t < s->n_thr; ++t) free(s->scratch[t]);
  free(s->scratch);
  free(s);
}

/* ---------------------------------------------------------------------------
 * Pair pass: ProcessPairInteraction + torque cross products
 * (src/interaction_manager.cpp:323-397, 448-498)
 * ------------------------------------------------------------------------- */
static void process_pair(orc_sys *s, int64_t c) {
  int32_t b1 = s->c1[c], b2 = s->c2[c];
  double c1v[3] = {0, 0, 0}, c2v[3] = {0, 0, 0};
  double force[3] = {s->cf[0][c], s->cf[1][c], s->cf[2][c]};
  if (b1 == b2) {
    s->err = -2; /* Logger::Error: self-interaction */
    return;
  }
  /* same mesh and sharing a site (Bond::HasNeighbor, src/bond.cpp:118-121) */
  if (s->bond_fil[b1] == s->bond_fil[b2] &&
      abs(s->bond_idx[b1] - s->bond_idx[b2]) == 1) {
    s->no_ix[c] = 1;
  } else {
    double s1[3] = {s->sp[0][b1], s->sp[1][b1], s->sp[2][b1]};
    double u1[3] = {s->u[0][b1], s->u[1][b1], s->u[2][b1]};
    double s2[3] = {s->sp[0][b2], s->sp[1][b2], s->sp[2][b2]};
    double u2[3] = {s->u[0][b2], s->u[1][b2], s->u[2][b2]};
    double dr[3], dr_mag2;
    sphero(s->p.h, s1, u1, s->len[b1], s2, u2, s->len[b2], dr, &dr_mag2, c1v,
           c2v);
    if (!(dr_mag2 > s->rcut2)) {
      if (pair_force(s->c12, s->c6, s->p.f_cutoff, dr, dr_mag2, c1v, c2v, force)) {
        /* stress_ij = -dr_i F_j; pote of the WCA branch, 0 for the
         * max-force branch (include/simcore/wca_potential.hpp:28-31,
         * include/simcore/max_force_potential.hpp:35-38) */
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) s->cst[3 * i + j][c] = -dr[i] * force[j];
        double pote = 0.0;
        if (!(dr_mag2 < 1e-12)) {
          double rmag = sqrt(dr_mag2);
          double rinv = 1.0 / (rmag);
          double rinv2 = rinv * rinv;
          double r6 = rinv2 * rinv2 * rinv2;
          pote = r6 * (s->c12 * r6 - s->c6) + s->p.wca_eps;
        }
        s->cpote[c] = pote;
      }
      for (int i = 0; i < 3; ++i) s->cf[i][c] = force[i];
    }
  }
  double t1[3], t2[3];
  cross3(c1v, force, t1);
  cross3(c2v, force, t2);
  for (int i = 0; i < 3; ++i) {
    s->ct1[i][c] = t1[i];
    s->ct2[i][c] = t2[i];
  }
}

/* ---------------------------------------------------------------------------
 * Filament::UpdatePosition (src/filament.cpp:394-404) for one filament.
 * ------------------------------------------------------------------------- */
static void update_filament(orc_sys *s, int f, int midstep, double *wk) {
  const int n = s->site_off[f + 1] - s->site_off[f];
  const int64_t s0 = s->site_off[f];
  const int64_t b0 = s0 - f;
  const double bl = s->bl[f];
  /* scratch layout */
  double *fs = wk;               /* 3n site forces        */
  double *tn = fs + 3 * n;       /* 3n tangents           */
  double *cs = tn + 3 * n;       /* n cos(theta)          */
  double *ga = cs + n;           /* 9n gamma inverse      */
  double *ta = ga + 9 * n;       /* n lower               */
  double *tb = ta + n;           /* n diag                */
  double *tc = tb + n;           /* n upper               */
  double *td = tc + n;           /* n rhs / tensions      */
  double *dt = td + n;           /* n+1 det_t             */
  double *db = dt + n + 1;       /* n+1 det_b             */
  double *ke = db + n + 1;       /* n k_eff               */
  double *uo = ke + n;           /* 3n new orientations   */
#define SU(i, d) (s->u[d][b0 + ((i) < n - 1 ? (i) : n - 2)]) /* site orient */
#define FS(i, d) fs[3 * (i) + (d)]
#define TN(i, d) tn[3 * (i) + (d)]
  memset(fs, 0, sizeof(double) * 3 * n); /* Mesh::ZeroForce */

  /* ApplyInteractionForces (src/filament.cpp:992-1048); driving is zero */
  {
    const double *U[3] = {s->u[0] + b0, s->u[1] + b0, s->u[2] + b0};
    const double *BF[3] = {s->bf[0] + b0, s->bf[1] + b0, s->bf[2] + b0};
    const double *BT[3] = {s->bt[0] + b0, s->bt[1] + b0, s->bt[2] + b0};
    bond_site_forces(n, bl, U, BF, BT, fs);
  }
  /* CalculateAngles (src/filament.cpp:430-437) */
  for (int i = 0; i < n - 2; ++i) {
    double a[3] = {SU(i, 0), SU(i, 1), SU(i, 2)};
    double b[3] = {SU(i + 1, 0), SU(i + 1, 1), SU(i + 1, 2)};
    cs[i] = dot3(a, b);
  }
  /* CalculateTangents -> Site::CalcTangent (src/site.cpp:88-102) */
  for (int i = 0; i < n; ++i) {
    if (i == 0 || i == n - 1) {
      int bb = (i == 0 ? 0 : n - 2);
      for (int d = 0; d < 3; ++d) TN(i, d) = s->u[d][b0 + bb];
    } else {
      double t[3];
      for (int d = 0; d < 3; ++d) t[d] = s->u[d][b0 + i - 1] + s->u[d][b0 + i];
      normalize3(t);
      for (int d = 0; d < 3; ++d) TN(i, d) = t[d];
    }
  }
  if (midstep) {
    if (s->p.stoch_flag) {
      /* ConstructUnprojectedRandomForces (src/filament.cpp:518-548) */
      const double sp_ = s->sig_perp[f], spa = s->sig_par[f];
      for (int i = 0; i < n; ++i) {
        double xi[3], xt[3];
        if (s->xi_buf) {
          const int64_t j = s->n_midsteps - s->xi_m0;
          for (int d = 0; d < 3; ++d)
            xi[d] = (j >= 0 && j < s->xi_nm) ? s->xi_buf[(j * s->ns + s0 + i) * 3 + d] : 0.0;
        } else {
          orc_noise(s->p.seed, s0 + i, s->n_midsteps, xi);
        }
        double t0 = TN(i, 0), t1 = TN(i, 1), t2 = TN(i, 2);
        xt[0] = t0 * t0 * xi[0] + t0 * t1 * xi[1] + t0 * t2 * xi[2];
        xt[1] = t1 * t1 * xi[1] + t0 * t1 * xi[0] + t1 * t2 * xi[2];
        xt[2] = t2 * t2 * xi[2] + t0 * t2 * xi[0] + t1 * t2 * xi[1];
        for (int d = 0; d < 3; ++d)
          s->fr[d][s0 + i] = sp_ * xi[d] + (spa - sp_) * xt[d];
      }
      /* GeometricallyProjectRandomForces (src/filament.cpp:550-599) */
      for (int i = 0; i < n - 1; ++i) {
        double tmp[3], u[3] = {SU(i, 0), SU(i, 1), SU(i, 2)};
        for (int d = 0; d < 3; ++d) tmp[d] = s->fr[d][s0 + i + 1] - s->fr[d][s0 + i];
        td[i] = dot3(tmp, u);
        tb[i] = 2;
        if (i > 0) {
          tc[i - 1] = -cs[i - 1];
          ta[i - 1] = -cs[i - 1];
        }
      }
      orc_tridiag(n - 1, ta, tb, tc, td);
      double fp[3];
      for (int d = 0; d < 3; ++d) fp[d] = s->fr[d][s0] + td[0] * SU(0, d);
      double fq[3];
      for (int d = 0; d < 3; ++d)
        fq[d] = s->fr[d][s0 + n - 1] - td[n - 2] * SU(n - 2, d);
      /* interior sites read their own unprojected value only */
      for (int i = 1; i < n - 1; ++i)
        for (int d = 0; d < 3; ++d)
          s->fr[d][s0 + i] = s->fr[d][s0 + i] + td[i] * SU(i, d) - td[i - 1] * SU(i - 1, d);
      for (int d = 0; d < 3; ++d) {
        s->fr[d][s0] = fp[d];
        s->fr[d][s0 + n - 1] = fq[d];
      }
    }
    /* Mesh::UpdatePrevPositions (src/mesh.cpp:175-182) */
    for (int d = 0; d < 3; ++d)
      for (int i = 0; i < n; ++i) s->rp[d][s0 + i] = s->r[d][s0 + i];
  }
  /* AddRandomForces (src/filament.cpp:601-606) */
  if (s->p.stoch_flag)
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) FS(i, d) += s->fr[d][s0 + i];

  /* CalculateBendingForces, 3D (src/filament.cpp:608-776) */
  {
    dt[0] = 1;
    dt[1] = 2;
    db[n] = 1;
    db[n - 1] = 2;
    for (int i = 2; i < n; ++i) {
      dt[i] = 2 * dt[i - 1] - (-cs[i - 2]) * (-cs[i - 2]) * dt[i - 2];
      db[n - i] = 2 * db[n - i + 1] - (-cs[n - i - 1]) * (-cs[n - i - 1]) * db[n - i + 2];
    }
    double det_g = dt[n - 1];
    for (int i = 0; i < n - 2; ++i) {
      double ginv = cs[i] * dt[i] * db[i + 3] / det_g;
      ke[i] = (s->kappa[f] + bl * ginv) / (bl * bl);
    }
    for (int k = 0; k < n; ++k) {
      double fsite[3] = {0.0, 0.0, 0.0};
      if (k > 1) {
        double u1[3] = {SU(k - 2, 0), SU(k - 2, 1), SU(k - 2, 2)};
        double u2[3] = {SU(k - 1, 0), SU(k - 1, 1), SU(k - 1, 2)};
        fsite[0] += ke[k - 2] * ((1 - u2[0] * u2[0]) * u1[0] - u2[0] * u2[1] * u1[1] - u2[0] * u2[2] * u1[2]);
        fsite[1] += ke[k - 2] * ((1 - u2[1] * u2[1]) * u1[1] - u2[1] * u2[0] * u1[0] - u2[1] * u2[2] * u1[2]);
        fsite[2] += ke[k - 2] * ((1 - u2[2] * u2[2]) * u1[2] - u2[2] * u2[0] * u1[0] - u2[2] * u2[1] * u1[1]);
      }
      if (k > 0 && k < n - 1) {
        double u1[3] = {SU(k - 1, 0), SU(k - 1, 1), SU(k - 1, 2)};
        double u2[3] = {SU(k, 0), SU(k, 1), SU(k, 2)};
        fsite[0] += ke[k - 1] * ((1 - u1[0] * u1[0]) * u2[0] - u1[0] * u1[1] * u2[1] - u1[0] * u1[2] * u2[2] -
                                 ((1 - u2[0] * u2[0]) * u1[0] - u2[0] * u2[1] * u1[1] - u2[0] * u2[2] * u1[2]));
        fsite[1] += ke[k - 1] * ((1 - u1[1] * u1[1]) * u2[1] - u1[1] * u1[0] * u2[0] - u1[1] * u1[2] * u2[2] -
                                 ((1 - u2[1] * u2[1]) * u1[1] - u2[1] * u2[0] * u1[0] - u2[1] * u2[2] * u1[2]));
        fsite[2] += ke[k - 1] * ((1 - u1[2] * u1[2]) * u2[2] - u1[2] * u1[0] * u2[0] - u1[2] * u1[1] * u2[1] -
                                 ((1 - u2[2] * u2[2]) * u1[2] - u2[2] * u2[0] * u1[0] - u2[1] * u2[2] * u1[1]));
      }
      if (k < n - 2) {
        double u1[3] = {SU(k, 0), SU(k, 1), SU(k, 2)};
        double u2[3] = {SU(k + 1, 0), SU(k + 1, 1), SU(k + 1, 2)};
        fsite[0] -= ke[k] * ((1 - u1[0] * u1[0]) * u2[0] - u1[0] * u1[1] * u2[1] - u1[0] * u1[2] * u2[2]);
        fsite[1] -= ke[k] * ((1 - u1[1] * u1[1]) * u2[1] - u1[1] * u1[0] * u2[0] - u1[1] * u1[2] * u2[2]);
        fsite[2] -= ke[k] * ((1 - u1[2] * u1[2]) * u2[2] - u1[2] * u1[0] * u2[0] - u1[2] * u1[1] * u2[1]);
      }
      for (int d = 0; d < 3; ++d) FS(k, d) += fsite[d];
    }
  }

  /* CalculateTensions (src/filament.cpp:778-838) */
  const double gpar = s->gpar[f], gperp = s->gperp[f];
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        ga[9 * i + 3 * a + b] = 1.0 / gpar * (TN(i, a) * TN(i, b)) +
                                1.0 / gperp * ((a == b ? 1 : 0) - TN(i, a) * TN(i, b));
  for (int i = 0; i < n - 1; ++i) {
    double fd[3];
    const double *g1 = ga + 9 * i, *g2 = ga + 9 * (i + 1);
    for (int a = 0; a < 3; ++a) {
      double ta_ = g1[3 * a] * FS(i, 0) + g1[3 * a + 1] * FS(i, 1);
      ta_ += g1[3 * a + 2] * FS(i, 2);
      double tb_ = g2[3 * a] * FS(i + 1, 0) + g2[3 * a + 1] * FS(i + 1, 1);
      tb_ += g2[3 * a + 2] * FS(i + 1, 2);
      fd[a] = tb_ - ta_;
    }
    double u2[3] = {SU(i, 0), SU(i, 1), SU(i, 2)};
    double t1[3] = {TN(i, 0), TN(i, 1), TN(i, 2)};
    double t2[3] = {TN(i + 1, 0), TN(i + 1, 1), TN(i + 1, 2)};
    td[i] = dot3(u2, fd);
    double a1 = dot3(t1, u2), a2 = dot3(t2, u2);
    tb[i] = 2.0 / gperp + (1.0 / gpar - 1.0 / gperp) * (a1 * a1 + a2 * a2);
    if (i > 0) {
      double u1[3] = {SU(i - 1, 0), SU(i - 1, 1), SU(i - 1, 2)};
      tc[i - 1] = -1.0 / gperp * dot3(u2, u1) -
                  (1.0 / gpar - 1.0 / gperp) * (dot3(t1, u1) * dot3(t1, u2));
      ta[i - 1] = tc[i - 1];
    }
  }
  orc_tridiag(n - 1, ta, tb, tc, td);
  for (int i = 0; i < n - 1; ++i) s->tension[b0 + i] = td[i];

  /* UpdateSitePositions (src/filament.cpp:840-927) */
  {
    double delta = (midstep ? 0.5 * s->p.delta : s->p.delta);
    for (int d = 0; d < 3; ++d) FS(0, d) += td[0] * SU(0, d);
    for (int d = 0; d < 3; ++d) FS(n - 1, d) += -td[n - 2] * SU(n - 2, d);
    for (int i = 1; i < n - 1; ++i)
      for (int d = 0; d < 3; ++d)
        FS(i, d) += td[i] * SU(i, d) - td[i - 1] * SU(i - 1, d);
    for (int i = 0; i < n; ++i) {
      const double *g = ga + 9 * i;
      for (int a = 0; a < 3; ++a) {
        double ft = g[3 * a] * FS(i, 0) + g[3 * a + 1] * FS(i, 1);
        ft += g[3 * a + 2] * FS(i, 2);
        s->r[a][s0 + i] = s->rp[a][s0 + i] + delta * ft;
      }
    }
    for (int i = 0; i < n - 1; ++i) {
      double rd[3], m = 0.0;
      for (int d = 0; d < 3; ++d) {
        rd[d] = s->r[d][s0 + i + 1] - s->r[d][s0 + i];
        m += rd[d] * rd[d];
      }
      m = sqrt(m);
      for (int d = 0; d < 3; ++d) uo[3 * i + d] = rd[d] / m;
    }
    for (int d = 0; d < 3; ++d) uo[3 * (n - 1) + d] = uo[3 * (n - 2) + d];
    /* CheckBondLengths (src/filament.cpp:934-951) */
    int renorm = 0;
    for (int i = 1; i < n; ++i) {
      double a = 0.0;
      for (int d = 0; d < 3; ++d) {
        double t = s->r[d][s0 + i] - s->r[d][s0 + i - 1];
        a += t * t;
      }
      a = sqrt(a);
      double err = O_ABS(bl - a) / bl;
      if (err > 1e-3) renorm = 1;
    }
    if (renorm) {
      s->norm_switch[f] = !s->norm_switch[f];
      s->n_normalize[f] = 0;
      if (s->norm_switch[f]) {
        for (int i = 1; i < n; ++i)
          for (int d = 0; d < 3; ++d)
            s->r[d][s0 + i] = s->r[d][s0 + i - 1] + bl * uo[3 * (i - 1) + d];
      } else {
        for (int i = n - 1; i > 0; --i)
          for (int d = 0; d < 3; ++d)
            s->r[d][s0 + i - 1] = s->r[d][s0 + i] - bl * uo[3 * (i - 1) + d];
      }
    }
    s->n_normalize[f]++;
  }
  /* Mesh::UpdateBondPositions (src/mesh.cpp:324-334) */
  for (int i = 0; i < n - 1; ++i) bond_reinit(s, b0 + i, s0 + i);
#undef SU
#undef FS
#undef TN
}

/* Mesh::UpdateDrTot (src/mesh.cpp:533-544) with Object::UpdateDrTot called
 * twice per site (once directly, once inside GetDrTot). */
static void update_dr_tot(orc_sys *s, int f) {
  for (int64_t i = s->site_off[f]; i < s->site_off[f + 1]; ++i) {
    for (int rep = 0; rep < 2; ++rep)
      for (int d = 0; d < 3; ++d) {
        double dr = s->r[d][i] - s->dr0[d][i];
        s->acc[i] += dr * dr;
      }
    if (s->acc[i] > s->dr_tot[f]) s->dr_tot[f] = s->acc[i];
  }
}

/* ---------------------------------------------------------------------------
 * Crosslinks
 * ------------------------------------------------------------------------- */
/* Anchor::UpdateAnchorPositionToBond (src/anchor.cpp:236-248) with the bond
 * state of Bond::ReInit: position, orientation and current length. */
static void anchor_point(const orc_sys *s, int32_t b, double lam, double *P, double *u,
                         double *len) {
  for (int i = 0; i < 3; ++i) u[i] = s->u[i][b];
  *len = s->len[b];
  for (int i = 0; i < 3; ++i) P[i] = s->pos[i][b] - (0.5 * *len - lam) * u[i];
}

static void scaled(const double *h_inv, const double *P, double *sp) {
  for (int i = 0; i < 3; ++i) {
    double si = 0.0;
    for (int j = 0; j < 3; ++j) si += h_inv[3 * i + j] * P[j];
    si -= O_NINT(si);
    sp[i] = si;
  }
}

double orc_xlink_force(const double *h, const double *h_inv, const double *p0,
                       const double *p1, double k, double r0, double kc, double *out) {
  double s0[3], s1[3], ds[3], dr[3], dr2 = 0.0;
  scaled(h_inv, p0, s0);
  scaled(h_inv, p1, s1);
  /* MinimumDistance::PointPoint (src/minimum_distance.cpp:71-101) */
  for (int i = 0; i < 3; ++i) {
    ds[i] = s1[i] - s0[i];
    ds[i] -= O_NINT(ds[i]);
  }
  for (int i = 0; i < 3; ++i) {
    dr[i] = 0.0;
    for (int j = 0; j < 3; ++j) dr[i] += h[3 * i + j] * ds[j];
  }
  for (int i = 0; i < 3; ++i) dr2 += dr[i] * dr[i];
  /* Crosslink::CalculateTetherForces (src/crosslink.cpp:246-273) */
  const double length = sqrt(dr2);
  const double stretch = length - r0;
  double tf;
  if (kc >= 0 && stretch < 0)
    tf = kc * stretch;
  else
    tf = k * stretch;
  for (int i = 0; i < 3; ++i) out[i] = tf * (dr[i] / length);
  return length;
}

/* even steps: forces from the current bonds, stored per crosslink */
static void xlink_forces(orc_sys *s) {
  for (int64_t x = 0; x < s->nx; ++x) {
    double P[2][3], u[2][3], len[2];
    for (int a = 0; a < 2; ++a)
      anchor_point(s, s->xb[2 * x + a], s->xlam[2 * x + a], P[a], u[a], &len[a]);
    double F[3];
    orc_xlink_force(s->p.h, s->p.h_inv, P[0], P[1], s->xp.k_spring, s->xp.rest_length,
                    s->xp.k_compress, F);
    double *o = s->xf + 9 * x;
    for (int i = 0; i < 3; ++i) o[i] = F[i];
    /* Anchor::ApplyAnchorForces (src/anchor.cpp:132-146): lever
     * (lambda - l/2) u; anchor 1 carries -F */
    for (int a = 0; a < 2; ++a) {
      double dl[3], Fa[3], T[3];
      for (int i = 0; i < 3; ++i) {
        dl[i] = (s->xlam[2 * x + a] - 0.5 * len[a]) * u[a][i];
        Fa[i] = a == 0 ? F[i] : -F[i];
      }
      cross3(dl, Fa, T);
      for (int i = 0; i < 3; ++i) o[3 + 3 * a + i] = T[i];
    }
  }
}

/* CrosslinkSpecies::ApplyCrosslinkTetherForces: serial, crosslink order */
static void xlink_apply(orc_sys *s) {
  for (int64_t x = 0; x < s->nx; ++x) {
    const double *o = s->xf + 9 * x;
    for (int a = 0; a < 2; ++a) {
      const int32_t b = s->xb[2 * x + a];
      for (int i = 0; i < 3; ++i) s->bf[i][b] += a == 0 ? o[i] : -o[i];
      for (int i = 0; i < 3; ++i) s->bt[i][b] += o[3 + 3 * a + i];
    }
  }
}

static double unit_draw(uint32_t v) { return ((double)v + 0.5) * 2.3283064365386962890625e-10; }

/* even steps after the forces: unbinding / rebinding */
static void xlink_events(orc_sys *s) {
  const int64_t m = s->i_step / 2;
  const uint32_t key[2] = {(uint32_t)s->xp.seed, (uint32_t)(s->xp.seed >> 32)};
  const int n = s->n1d, R = s->xreach;
  const double rb2 = s->xp.r_bind * s->xp.r_bind;
  for (int64_t x = 0; x < s->nx; ++x) {
    uint32_t ctr[4] = {(uint32_t)x, (uint32_t)((uint64_t)x >> 32), (uint32_t)m, 1u}, o[4];
    orc_philox4x32(ctr, key, o);
    if (!(unit_draw(o[0]) < s->xpoff)) continue;
    const int am = unit_draw(o[1]) < 0.5 ? 0 : 1, as = 1 - am;
    const int32_t bs = s->xb[2 * x + as];
    double P[3], u[3], len, sP[3];
    anchor_point(s, bs, s->xlam[2 * x + as], P, u, &len);
    scaled(s->p.h_inv, P, sP);
    const int32_t key0 = orc_cell_key(n, sP);
    const int cx = key0 / (n * n), cy = (key0 / n) % n, cz = key0 % n;
    const int span = 2 * R + 1 >= n ? n : 2 * R + 1;
    const int off0 = 2 * R + 1 >= n ? 0 : -R;
    int32_t best = -1;
    uint32_t best_score = 0;
    for (int dx = 0; dx < span; ++dx)
      for (int dy = 0; dy < span; ++dy)
        for (int dz = 0; dz < span; ++dz) {
          const int X = ((cx + off0 + dx) % n + n) % n, Y = ((cy + off0 + dy) % n + n) % n;
          const int Z = ((cz + off0 + dz) % n + n) % n;
          const int32_t c = ((int32_t)X * n + Y) * n + Z;
          for (int32_t j = s->cell_start[c]; j < s->cell_start[c + 1]; ++j) {
            const int32_t b = s->cell_obj[j];
            if (b == bs) continue;
            double ds[3], dr[3], t = 0.0, d2 = 0.0;
            for (int i = 0; i < 3; ++i) {
              ds[i] = s->sp[i][b] - sP[i];
              ds[i] -= O_NINT(ds[i]);
            }
            for (int i = 0; i < 3; ++i) {
              dr[i] = 0.0;
              for (int k = 0; k < 3; ++k) dr[i] += s->p.h[3 * i + k] * ds[k];
            }
            for (int i = 0; i < 3; ++i) t += dr[i] * s->u[i][b];
            const double hl = 0.5 * s->len[b];
            if (t > hl) t = hl;
            if (t < -hl) t = -hl;
            for (int i = 0; i < 3; ++i) {
              const double e = dr[i] - t * s->u[i][b];
              d2 += e * e;
            }
            if (!(d2 <= rb2)) continue;
            uint32_t c2[4] = {(uint32_t)x, (uint32_t)m, (uint32_t)b, 2u}, q[4];
            orc_philox4x32(c2, key, q);
            if (best < 0 || q[0] < best_score || (q[0] == best_score && b < best)) {
              best = b;
              best_score = q[0];
            }
          }
        }
    if (best < 0) continue;
    s->xb[2 * x + am] = best;
    s->xlam[2 * x + am] = unit_draw(o[2]) * s->bl[s->bond_fil[best]];
    for (int i = 0; i < 9; ++i) s->xf[9 * x + i] = 0.0;
    s->xl_events++;
  }
}

int orc_load_crosslinks(orc_sys *s, int64_t nx, const int32_t *bonds, const double *lambda,
                        const orc_xlink_params *p) {
  if (!s || nx < 0 || (nx > 0 && (!bonds || !lambda)) || !p) return -1;
  for (int64_t i = 0; i < 2 * nx; ++i)
    if (bonds[i] < 0 || bonds[i] >= s->nb) return -1;
  free(s->xb);
  free(s->xlam);
  free(s->xf);
  s->nx = nx;
  s->xb = xcalloc((size_t)(2 * nx), sizeof(int32_t));
  s->xlam = xcalloc((size_t)(2 * nx), sizeof(double));
  s->xf = xcalloc((size_t)(9 * nx), sizeof(double));
  memcpy(s->xb, bonds, sizeof(int32_t) * (size_t)(2 * nx));
  memcpy(s->xlam, lambda, sizeof(double) * (size_t)(2 * nx));
  s->xp = *p;
  s->xpoff = 1.0 - exp(-p->k_off * s->p.delta);
  double lmax = 0.0;
  for (int f = 0; f < s->nf; ++f)
    if (s->bl[f] > lmax) lmax = s->bl[f];
  /* cells to search: the range plus half a bond, plus one cell for
   * assignments made at the last rebuild */
  s->xreach = (int)ceil((p->r_bind + 0.5 * lmax) / s->cell_len) + 1;
  s->xl_events = 0;
  return 0;
}

void orc_get_crosslinks(const orc_sys *s, int32_t *bonds, double *lambda, double *force) {
  if (bonds) memcpy(bonds, s->xb, sizeof(int32_t) * (size_t)(2 * s->nx));
  if (lambda) memcpy(lambda, s->xlam, sizeof(double) * (size_t)(2 * s->nx));
  if (force) memcpy(force, s->xf, sizeof(double) * (size_t)(9 * s->nx));
}

int64_t orc_xlink_events(const orc_sys *s) { return s->xl_events; }

int orc_step(orc_sys *s, int nsteps) {
  for (int step = 0; step < nsteps; ++step) {
    s->i_step++;
    /* Simulation::ZeroForces (bond F/T; site forces start at zero inside
     * update_filament) */
    for (int d = 0; d < 3; ++d) {
      memset(s->bf[d], 0, sizeof(double) * (size_t)s->nb);
      memset(s->bt[d], 0, sizeof(double) * (size_t)s->nb);
    }
    /* InteractionManager::Interact -> CrosslinkManager::UpdateCrosslinks
     * (src/interaction_manager.cpp:70, src/crosslink_species.cpp:218-238) */
    if (s->nx > 0) {
      if (s->i_step % 2 == 0) xlink_forces(s);
      xlink_apply(s);
      if (s->i_step % 2 == 0 && s->xpoff > 0) xlink_events(s);
    }
    /* CheckUpdateInteractions -> GetDrMax (src/interaction_manager.cpp:
     * 284-312, include/simcore/species.hpp:105-120) */
    if (!s->last_midstep) {
#pragma omp parallel for schedule(static) num_threads(s->n_thr)
      for (int f = 0; f < s->nf; ++f) update_dr_tot(s, f);
    }
    double max_dr = 0;
    for (int f = 0; f < s->nf; ++f)
      if (s->dr_tot[f] > max_dr) max_dr = s->dr_tot[f];
    if (max_dr > 10000) return -1; /* throw std::runtime_error */
    if (max_dr > s->dr_update) rebuild(s);
    /* CalculatePairInteractions: OpenMP static chunks */
    s->err = 0;
    int64_t nc = s->nc;
#pragma omp parallel for schedule(static) num_threads(s->n_thr)
    for (int64_t c = 0; c < nc; ++c) process_pair(s, c);
    if (s->err) return s->err;
    if (s->i_update == 0) {
      /* order-preserving cull of no_interaction pairs */
      int64_t w = 0;
      for (int64_t c = 0; c < s->nc; ++c) {
        if (s->no_ix[c]) continue;
        s->c1[w] = s->c1[c];
        s->c2[w] = s->c2[c];
        s->no_ix[w] = 0;
        for (int d = 0; d < 3; ++d) {
          s->cf[d][w] = s->cf[d][c];
          s->ct1[d][w] = s->ct1[d][c];
          s->ct2[d][w] = s->ct2[d][c];
        }
        for (int k = 0; k < 9; ++k) s->cst[k][w] = s->cst[k][c];
        s->cpote[w] = s->cpote[c];
        w++;
      }
      s->nc = w;
    }
    /* ApplyPairInteractions: serial scatter (src/interaction_manager.cpp:
     * 500-517) */
    memset(s->bpote, 0, sizeof(double) * (size_t)s->nb);
    for (int64_t c = 0; c < s->nc; ++c) {
      int32_t b1 = s->c1[c], b2 = s->c2[c];
      for (int d = 0; d < 3; ++d) s->bf[d][b1] += s->cf[d][c];
      for (int d = 0; d < 3; ++d) s->bf[d][b2] -= s->cf[d][c];
      for (int d = 0; d < 3; ++d) s->bt[d][b1] += s->ct1[d][c];
      for (int d = 0; d < 3; ++d) s->bt[d][b2] -= s->ct2[d][c];
      s->bpote[b1] += s->cpote[c];
      s->bpote[b2] += s->cpote[c];
      for (int k = 0; k < 9; ++k) s->stress_acc[k] += s->cst[k][c];
    }
    s->i_update++;
    /* FilamentSpecies::UpdatePositions (src/filament_species.cpp:59-85) */
    int mid = s->midstep;
#pragma omp parallel num_threads(s->n_thr)
    {
#ifdef _OPENMP
      double *wk = s->scratch[omp_get_thread_num()];
#else
      double *wk = s->scratch[0];
#endif
#pragma omp for schedule(static)
      for (int f = 0; f < s->nf; ++f) update_filament(s, f, mid, wk);
    }
    s->last_midstep = mid;
    if (mid) s->n_midsteps++;
    s->midstep = !s->midstep;
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * Taps
 * ------------------------------------------------------------------------- */
int64_t orc_n_sites(const orc_sys *s) { return s->ns; }
int64_t orc_n_bonds(const orc_sys *s) { return s->nb; }
int64_t orc_n_candidates(const orc_sys *s) { return s->nc; }
int64_t orc_rebuild_count(const orc_sys *s) { return s->rebuilds; }
int64_t orc_step_index(const orc_sys *s) { return s->i_step; }
int orc_n_cells_1d(const orc_sys *s) { return s->n1d; }
double orc_cell_length(const orc_sys *s) { return s->cell_len; }

void orc_get_candidates(const orc_sys *s, int32_t *b1, int32_t *b2) {
  memcpy(b1, s->c1, sizeof(int32_t) * (size_t)s->nc);
  memcpy(b2, s->c2, sizeof(int32_t) * (size_t)s->nc);
}
void orc_get_slot_forces(const orc_sys *s, double *f) {
  for (int d = 0; d < 3; ++d)
    memcpy(f + d * s->nc, s->cf[d], sizeof(double) * (size_t)s->nc);
}
void orc_get_cell_keys(const orc_sys *s, int32_t *keys) {
  memcpy(keys, s->key, sizeof(int32_t) * (size_t)s->nb);
}
static void get3(double *dst, double *const src[3], int64_t n) {
  for (int d = 0; d < 3; ++d) memcpy(dst + d * n, src[d], sizeof(double) * (size_t)n);
}
void orc_get_sites(const orc_sys *s, double *r, double *r_prev, double *f_rand) {
  if (r) get3(r, s->r, s->ns);
  if (r_prev) get3(r_prev, s->rp, s->ns);
  if (f_rand) get3(f_rand, s->fr, s->ns);
}
void orc_get_bonds(const orc_sys *s, double *pos, double *spos, double *u,
                   double *len) {
  if (pos) get3(pos, s->pos, s->nb);
  if (spos) get3(spos, s->sp, s->nb);
  if (u) get3(u, s->u, s->nb);
  if (len) memcpy(len, s->len, sizeof(double) * (size_t)s->nb);
}
void orc_get_bond_forces(const orc_sys *s, double *f, double *t) {
  if (f) get3(f, s->bf, s->nb);
  if (t) get3(t, s->bt, s->nb);
}
void orc_get_tensions(const orc_sys *s, double *tension) {
  memcpy(tension, s->tension, sizeof(double) * (size_t)s->nb);
}
void orc_get_dr(const orc_sys *s, double *acc, double *dr_max_per_fil) {
  if (acc) memcpy(acc, s->acc, sizeof(double) * (size_t)s->ns);
  if (dr_max_per_fil) memcpy(dr_max_per_fil, s->dr_tot, sizeof(double) * (size_t)s->nf);
}
/* The stress_ sum since the last call (InteractionManager::
 * ApplyPairInteractions adds every pair's stress every step,
 * src/interaction_manager.cpp:500-517; CalculatePressure reads and resets
 * it, :534-559), and each bond's potential energy of the last step. */
void orc_take_stress(orc_sys *s, double *out) {
  for (int k = 0; k < 9; ++k) {
    out[k] = s->stress_acc[k];
    s->stress_acc[k] = 0.0;
  }
}
void orc_get_bond_energies(const orc_sys *s, double *out) {
  memcpy(out, s->bpote, sizeof(double) * (size_t)s->nb);
}

/* Injects a pre-drawn noise stream: xi for the next n_mid midsteps, 3 per
 * site in site order (the draws RandomUniform() - 0.5 of
 * src/filament.cpp:527-547); n_mid = 0 returns to the counter stream. */
void orc_set_noise_buffer(orc_sys *s, const double *xi, int64_t n_mid) {
  free(s->xi_buf);
  s->xi_buf = NULL;
  s->xi_nm = 0;
  if (n_mid <= 0 || !xi) return;
  const size_t n = (size_t)n_mid * (size_t)s->ns * 3;
  s->xi_buf = malloc(sizeof(double) * n);
  memcpy(s->xi_buf, xi, sizeof(double) * n);
  s->xi_m0 = s->n_midsteps;
  s->xi_nm = n_mid;
}

void orc_set_sites(orc_sys *s, const double *r) {
  /* positions replaced from outside: bond state follows (Mesh::
   * UpdateBondPositions, src/mesh.cpp:324-334) */
  for (int d = 0; d < 3; ++d) memcpy(s->r[d], r + d * s->ns, sizeof(double) * (size_t)s->ns);
  for (int f = 0; f < s->nf; ++f)
    for (int i = 0; i < s->site_off[f + 1] - s->site_off[f] - 1; ++i)
      bond_reinit(s, s->site_off[f] - f + i, s->site_off[f] + i);
}

void orc_bending_forces(int n, const double *u, double kappa, double bl,
                        double *f_site, double *k_eff) {
  /* wraps the same code path as update_filament's bending block through a
   * single-filament temporary system is heavy; restate directly. */
  int nb = n - 1;
#define UU(i, d) u[(d)*nb + ((i) < n - 1 ? (i) : n - 2)]
  double *cs = calloc((size_t)n, sizeof(double));
  double *dt = calloc((size_t)n + 1, sizeof(double));
  double *db = calloc((size_t)n + 1, sizeof(double));
  for (int i = 0; i < n - 2; ++i) {
    double a[3] = {UU(i, 0), UU(i, 1), UU(i, 2)};
    double b[3] = {UU(i + 1, 0), UU(i + 1, 1), UU(i + 1, 2)};
    cs[i] = dot3(a, b);
  }
  dt[0] = 1;
  dt[1] = 2;
  db[n] = 1;
  db[n - 1] = 2;
  for (int i = 2; i < n; ++i) {
    dt[i] = 2 * dt[i - 1] - (-cs[i - 2]) * (-cs[i - 2]) * dt[i - 2];
    db[n - i] = 2 * db[n - i + 1] - (-cs[n - i - 1]) * (-cs[n - i - 1]) * db[n - i + 2];
  }
  double det_g = dt[n - 1];
  for (int i = 0; i < n - 2; ++i) {
    double ginv = cs[i] * dt[i] * db[i + 3] / det_g;
    k_eff[i] = (kappa + bl * ginv) / (bl * bl);
  }
  for (int k = 0; k < n; ++k) {
    double fs[3] = {0.0, 0.0, 0.0};
    if (k > 1) {
      double u1[3] = {UU(k - 2, 0), UU(k - 2, 1), UU(k - 2, 2)};
      double u2[3] = {UU(k - 1, 0), UU(k - 1, 1), UU(k - 1, 2)};
      for (int d = 0; d < 3; ++d) {
        int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
        fs[d] += k_eff[k - 2] * ((1 - u2[d] * u2[d]) * u1[d] - u2[d] * u2[p] * u1[p] - u2[d] * u2[q] * u1[q]);
      }
    }
    if (k > 0 && k < n - 1) {
      double u1[3] = {UU(k - 1, 0), UU(k - 1, 1), UU(k - 1, 2)};
      double u2[3] = {UU(k, 0), UU(k, 1), UU(k, 2)};
      for (int d = 0; d < 3; ++d) {
        int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
        fs[d] += k_eff[k - 1] * ((1 - u1[d] * u1[d]) * u2[d] - u1[d] * u1[p] * u2[p] - u1[d] * u1[q] * u2[q] -
                                 ((1 - u2[d] * u2[d]) * u1[d] - u2[d] * u2[p] * u1[p] - u2[d] * u2[q] * u1[q]));
      }
    }
    if (k < n - 2) {
      double u1[3] = {UU(k, 0), UU(k, 1), UU(k, 2)};
      double u2[3] = {UU(k + 1, 0), UU(k + 1, 1), UU(k + 1, 2)};
      for (int d = 0; d < 3; ++d) {
        int p = d == 0 ? 1 : 0, q = d == 2 ? 1 : 2;
        fs[d] -= k_eff[k] * ((1 - u1[d] * u1[d]) * u2[d] - u1[d] * u1[p] * u2[p] - u1[d] * u1[q] * u2[q]);
      }
    }
    for (int d = 0; d < 3; ++d) f_site[d * n + k] = fs[d];
  }
#undef UU
  free(cs);
  free(dt);
  free(db);
}
