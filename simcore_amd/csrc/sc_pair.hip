// sc_pair.hip -- pair pass and force scatter, fused and bond-centric.
//
// Replaces InteractionManager::CalculatePairInteractions +
// ApplyPairInteractions (src/interaction_manager.cpp:448-517).  One thread
// owns one bond and walks the bond's incidence list (its candidates in
// list order, CSR built at rebuild time).  For every candidate it computes
// the pair with the candidate's own (obj1, obj2) ordering -- the thread of
// obj1 and the thread of obj2 run the same instruction sequence on the same
// inputs and obtain the same bits -- updates the sticky force slot when the
// pair is inside the cutoff, and adds +F, contact1 x F (obj1) or -F,
// -contact2 x F (obj2) in candidate order: the serial scatter's summation
// order, with no atomics and no second pass.
//
// Bond state is one record per bond (scaled position, orientation,
// length): 64 bytes of FP64 on the FP64 path, so each gather touches half a
// cache line; 32 bytes on the mixed path (32-bit fixed-point offset inside
// the bond's cell and the packed cell index, octahedral orientation, length
// as two FP32 values), decoded to a position in cell units in FP64.
//
// FP64 path: bit-for-bit the reference's arithmetic.  Mixed path: the same
// FP64 geometry and branch rules with the divisions and square roots taken
// from the hardware FP32 reciprocal / reciprocal square root (the latter
// refined in FP64), and FP32 force slots.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

namespace {

// A bond as the pair pass uses it.  FP64 path: the scaled position, as
// Bond::ReInit stores it.  Mixed path: the position in cell units, cell
// index plus the record's 32-bit offset (exact in FP64), so that the
// separation of two bonds is an exact difference and one multiplication.
struct BondRec {
  double s[3], u[3], len;
};

template <bool MIXED>
__device__ __forceinline__ BondRec load_rec(const double *br, const Geo &g, int64_t b) {
  BondRec r;
  if (MIXED) {
#pragma clang fp contract(fast)
    const uint4 *rec = reinterpret_cast<const uint4 *>(br + 4 * b);
    const uint4 a = rec[0], v = rec[1];
    const int packed = (int)v.z;
    constexpr double k32 = 1.0 / 4294967296.0;
    r.s[0] = (double)(packed & 1023) + (double)a.x * k32;
    r.s[1] = (double)((packed >> 10) & 1023) + (double)a.y * k32;
    r.s[2] = (double)((packed >> 20) & 1023) + (double)a.z * k32;
    oct_decode(v.x, v.y, r.u);
    r.len = (double)__uint_as_float(a.w) + (double)__uint_as_float(v.w);
  } else {
    rec_read_full(br, b, r.s, r.u, &r.len);
  }
  return r;
}

// Mixed path: separation r2 - r1 of two cell-unit positions with the
// minimum-image rule of the reference's ds -= NINT(ds) (|ds| >= half the
// box wraps), times the cell length.
__device__ __forceinline__ void sep_cells(const Geo &g, const double *p1, const double *p2,
                                          double *sep) {
#pragma clang fp contract(fast)
  const double n = (double)g.n1d, half = 0.5 * (double)g.n1d;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double ds = p2[k] - p1[k];
    const double w = ds >= half ? ds - n : (ds <= -half ? ds + n : ds);
    sep[k] = w * (g.h[4 * k] * g.inv_n1d);
  }
}

// Mixed-precision MinimumDistance::Sphero (src/minimum_distance.cpp:231-343):
// the reference's FP64 arithmetic and branch rules, except that the division
// by 1-(u1.u2)^2 multiplies by the hardware FP32 reciprocal (relative error
// ~1e-7).  A clamped line parameter is recomputed exactly from the other
// one, so the reciprocal's error reaches only unclamped parameters
// (|lambda| <= l/2).  On CDNA4 an FP64 FMA issues at the rate of an
// unpacked FP32 one, so keeping the geometry in FP64 costs nothing and
// needs no separate near-parallel path; only the divisions and square roots
// (many instructions in IEEE FP64) are replaced.  Multiply-adds are fused
// (one rounding each); both owners of a pair still run the same
// instructions on the same inputs.
__device__ __forceinline__ void sphero_mixed(const double *dr, const double *u1, double hl1,
                                             const double *u2, double hl2, double *rmin,
                                             double *rm2, double *lam_o, double *mu_o) {
#pragma clang fp contract(fast)
  double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    a += dr[i] * u1[i];
    b += dr[i] * u2[i];
    c += u1[i] * u2[i];
  }
  double lambda, mu;
  const double denom = 1.0 - c * c;
  if (denom < 1.0e-12) {
    lambda = 0.5 * a;
    mu = -0.5 * b;
  } else {
    const double id = (double)__builtin_amdgcn_rcpf((float)denom);
    lambda = (a - c * b) * id;
    mu = (-b + c * a) * id;
  }
  const double lm = dabs(lambda), mm = dabs(mu);
  if (lm > hl1 && mm > hl2) {
    const double la = dsign(hl1, lambda);
    double ma = -b + la * c;
    if (dabs(ma) > hl2) ma = dsign(hl2, ma);
    const double mb = dsign(hl2, mu);
    double lb = a + mb * c;
    if (dabs(lb) > hl1) lb = dsign(hl1, lb);
    double m2a = 0.0, m2b = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double ra = dr[i] - la * u1[i] + ma * u2[i];
      const double rb = dr[i] - lb * u1[i] + mb * u2[i];
      m2a += ra * ra;
      m2b += rb * rb;
    }
    if (m2a < m2b) {
      lambda = la;
      mu = ma;
    } else {
      lambda = lb;
      mu = mb;
    }
  } else if (lm > hl1) {
    lambda = dsign(hl1, lambda);
    mu = -b + lambda * c;
    if (dabs(mu) > hl2) mu = dsign(hl2, mu);
  } else if (mm > hl2) {
    mu = dsign(hl2, mu);
    lambda = a + mu * c;
    if (dabs(lambda) > hl1) lambda = dsign(hl1, lambda);
  }
  double m2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rmin[i] = dr[i] - lambda * u1[i] + mu * u2[i];
    m2 += rmin[i] * rmin[i];
  }
  *rm2 = m2;
  *lam_o = lambda;
  *mu_o = mu;
}

// WCA force of WCAPotential::CalcPotential (include/simcore/wca_potential.hpp:
// 14-32) with 1/r from the hardware FP32 reciprocal square root refined by
// one FP64 Newton step (relative error ~1e-14), rounded to FP32 for the
// slot.  dr^2 >= 1e-12 here.
__device__ __forceinline__ float4 wca_mixed(const Geo &g, const double *rmin, double rm2) {
#pragma clang fp contract(fast)
  double rinv = (double)__builtin_amdgcn_rsqf((float)rm2);
  rinv = rinv * (1.5 - 0.5 * rm2 * rinv * rinv);
  const double rinv2 = rinv * rinv;
  const double r6 = rinv2 * rinv2 * rinv2;
  double ffac = -(12.0 * g.c12 * r6 - 6.0 * g.c6) * r6 * rinv;
  if (dabs(ffac) > g.fcut) ffac = (ffac >= 0.0 ? 1.0 : -1.0) * g.fcut;
  const double sc = ffac * rinv;
  return make_float4((float)(sc * rmin[0]), (float)(sc * rmin[1]), (float)(sc * rmin[2]),
                     (float)(rm2 * rinv));
}

}  // namespace

// Block index -> bond range.  Workgroups are dispatched to the 8 XCDs
// round-robin; with `xcd` set, the blocks one XCD receives cover one
// contiguous range of bonds, so with spatially ordered storage a bond's
// partners tend to be in the same XCD's L2.
__device__ __forceinline__ int64_t logical_block(int xcd) {
  const int64_t b = blockIdx.x;
  if (!xcd) return b;
  const int64_t n = gridDim.x, q = n / 8, r = n % 8;
  const int64_t x = b % 8, k = b / 8;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// The pair of incidence q {partner, role | written << 1 | owner << 2},
// computed with the candidate's own (obj1, obj2) orientation
// (src/interaction_manager.cpp:448-498): the incidence's slot copy updated
// when in range (and marked written), else the copy's force, 0 if it was
// not written since the rebuild (the reference's fresh pair record).  Both
// owners of a candidate run the same instructions on the same inputs and
// obtain the same bits, so their slot copies stay equal.  The slot's fourth
// value is the pair's signed distance (pair_force64): with the force it
// gives the pair's stress and potential energy, so the sticky record needs
// nothing else.  Returns the candidate's force in fc, the owner's contact
// parameter (lambda for obj1, mu for obj2) in *sc and the signed distance
// in *rs (0 for a slot not written since the rebuild).
template <bool MIXED>
__device__ __forceinline__ void pair_slot(const Dev &d, const Geo &g, const BondRec &own,
                                          int32_t q, int2 ent, const BondRec &oth, double *fc,
                                          double *sc, double *rs_out) {
  const bool second = (ent.y & 1) != 0;  // this bond is obj2 of the candidate
  const bool written = (ent.y & 2) != 0;
  // obj1 and obj2 by selection (no branch on the role)
  double u1[3], u2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    u1[k] = second ? oth.u[k] : own.u[k];
    u2[k] = second ? own.u[k] : oth.u[k];
  }
  const double l1 = second ? oth.len : own.len;
  const double l2 = second ? own.len : oth.len;
  double lam, mu, rs = 0.0;
  fc[0] = fc[1] = fc[2] = 0.0;
  bool wrote = false;
  if (!MIXED) {
    double s1[3], s2[3], dr[3], dr2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s1[k] = second ? oth.s[k] : own.s[k];
      s2[k] = second ? own.s[k] : oth.s[k];
    }
    sphero64(g.h, s1, u1, l1, s2, u2, l2, dr, &dr2, &lam, &mu);
    if (!(dr2 > g.rcut2)) {
      double f[3];
      if (pair_force64(g.c12, g.c6, g.fcut, dr, dr2, lam, mu, u1, u2, f, &rs)) {
        fc[0] = f[0];
        fc[1] = f[1];
        fc[2] = f[2];
        d.islot64[q] = make_double4(f[0], f[1], f[2], rs);
        wrote = true;
      }
    }
    if (!wrote && written) {
      const double4 o = d.islot64[q];
      fc[0] = o.x;
      fc[1] = o.y;
      fc[2] = o.z;
      rs = o.w;
    }
  } else {
    // obj2 - obj1: the owner-to-partner separation, negated for obj2 (the
    // minimum-image rule is odd, so both owners get the same bits)
    double sep[3], dr[3], dr2;
    sep_cells(g, own.s, oth.s, sep);
#pragma unroll
    for (int k = 0; k < 3; ++k) sep[k] = second ? -sep[k] : sep[k];
    sphero_mixed(sep, u1, 0.5 * l1, u2, 0.5 * l2, dr, &dr2, &lam, &mu);
    if (!(dr2 > g.rcut2)) {
      float4 f4;
      if (dr2 < 1.0e-12) {
        // MaxForcePotential branch (include/simcore/max_force_potential.hpp:
        // 11-40), FP64 as the reference; practically never taken
        double f[3];
        wrote = pair_force64(g.c12, g.c6, g.fcut, dr, dr2, lam, mu, u1, u2, f, &rs) != 0;
        f4 = make_float4((float)f[0], (float)f[1], (float)f[2], (float)rs);
      } else {
        f4 = wca_mixed(g, dr, dr2);
        wrote = true;
      }
      if (wrote) {
        d.islot32[q] = f4;
        fc[0] = f4.x;
        fc[1] = f4.y;
        fc[2] = f4.z;
        rs = f4.w;
      }
    }
    if (!wrote && written) {
      const float4 o = d.islot32[q];
      fc[0] = o.x;
      fc[1] = o.y;
      fc[2] = o.z;
      rs = o.w;
    }
  }
  *rs_out = rs;
  *sc = second ? mu : lam;
  if (wrote && !written) reinterpret_cast<int32_t *>(d.inc)[2 * (int64_t)q + 1] = ent.y | 2;
}

// The owner's share of a pair (src/interaction_manager.cpp:469-470,
// 500-517; src/minimum_distance.cpp:339-342): +F and contact1 x F for obj1,
// -F and -(contact2 x F) for obj2.  With f the signed force (-F for obj2)
// the torque is contact x f, the same bits as -(contact x F).  The mixed
// path fuses the cross product's multiply-adds.
template <bool MIXED>
__device__ __forceinline__ void pair_share(const double *u, double sc, const double *f,
                                           double *t) {
  double con[3];
  for (int k = 0; k < 3; ++k) con[k] = sc * u[k];
  if (MIXED) {
#pragma clang fp contract(fast)
    t[0] = con[1] * f[2] - con[2] * f[1];
    t[1] = con[2] * f[0] - con[0] * f[2];
    t[2] = con[0] * f[1] - con[1] * f[0];
  } else {
    cross3(con, f, t);
  }
}

// Per-bond sums, accumulated in FP64; FP32 storage on the mixed path (one
// rounding of the sum).
template <bool MIXED>
__device__ __forceinline__ void store_bond_sums(const Dev &d, int64_t nb, int64_t b,
                                                const double *F, const double *T) {
  for (int k = 0; k < 3; ++k) {
    if (MIXED) {
      d.bf32[k * nb + b] = (float)F[k];
      d.bt32[k * nb + b] = (float)T[k];
    } else {
      d.bf[k * nb + b] = F[k];
      d.bt[k * nb + b] = T[k];
    }
  }
}

// Stress of one candidate, xx yy zz xy xz yz: -dr_i F_j of
// WCAPotential::CalcPotential (include/simcore/wca_potential.hpp:28-30),
// from the force and the signed distance (dr = -(rs/|F|) F).
__device__ __forceinline__ void add_stress(const double *f, double rs, double *S) {
  const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
  if (!(f2 > 0.0)) return;
  const double k = rs / sqrt(f2);
  S[0] += k * f[0] * f[0];
  S[1] += k * f[1] * f[1];
  S[2] += k * f[2] * f[2];
  S[3] += k * f[0] * f[1];
  S[4] += k * f[0] * f[2];
  S[5] += k * f[1] * f[2];
}

// A block's stress sum, added to its slot of the accumulator (each block
// owns one slot, so no atomics); blocks of several wavefronts combine their
// wavefronts' sums through LDS.
__device__ __forceinline__ void block_stress(const Dev &d, double *S) {
  __shared__ double part[8][6];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k)
    for (int o = 32; o > 0; o >>= 1) S[k] += __shfl_xor(S[k], o, 64);
  if (lane == 0)
    for (int k = 0; k < 6; ++k) part[w][k] = S[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += part[i][threadIdx.x];
    d.stress_acc[6 * (int64_t)blockIdx.x + threadIdx.x] += t;
  }
}

template <bool MIXED, bool THERMO>
__global__ __launch_bounds__(512) void k_pair_fused(Dev d, Geo g, int xcd, int64_t blk0) {
  if (d.ctl->abort) return;
  const int64_t nb = g.nb;
  const int64_t b = (blk0 + logical_block(xcd)) * blockDim.x + threadIdx.x;
  double S[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (b < nb) {
    const BondRec own = load_rec<MIXED>(d.br, g, b);
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0}, E = 0.0;
    if (g.nx > 0) {
      // crosslink forces first, in crosslink order (the reference applies
      // them inside Interact before the pair forces,
      // src/interaction_manager.cpp:70, src/crosslink_species.cpp:262-266)
      const int32_t x0 = d.xinc_start[b], x1 = d.xinc_start[b + 1];
      for (int32_t q = x0; q < x1; ++q) {
        const int32_t e = d.xinc[q];
        const double *o = d.xf + 9 * (int64_t)(e >> 1);
        const int a = e & 1;
        for (int k = 0; k < 3; ++k) F[k] += a == 0 ? o[k] : -o[k];
        for (int k = 0; k < 3; ++k) T[k] += o[3 + 3 * a + k];
      }
    }
    const int32_t q0 = d.inc_start[b], q1 = d.inc_start[b + 1];
    // software pipeline: the next incidence's entry and partner record are
    // in flight while the current pair is computed
    int2 ent_n = make_int2(0, 0);
    BondRec oth_n{};
    if (q0 < q1) {
      ent_n = d.inc[q0];
      oth_n = load_rec<MIXED>(d.br, g, ent_n.x);
    }
    for (int32_t q = q0; q < q1; ++q) {
      const int2 ent = ent_n;
      const BondRec oth = oth_n;
      if (q + 1 < q1) {
        ent_n = d.inc[q + 1];
        oth_n = load_rec<MIXED>(d.br, g, ent_n.x);
      }
      double fc[3], sc, rs, cf[3], ct[3];
      pair_slot<MIXED>(d, g, own, q, ent, oth, fc, &sc, &rs);
      const double sg = (ent.y & 1) ? -1.0 : 1.0;
      for (int k = 0; k < 3; ++k) cf[k] = sg * fc[k];
      pair_share<MIXED>(own.u, sc, cf, ct);
      for (int k = 0; k < 3; ++k) F[k] += cf[k];
      for (int k = 0; k < 3; ++k) T[k] += ct[k];
      if (THERMO) {
        // Object::AddPotential on both bonds; the candidate's stress once
        // (src/interaction_manager.cpp:500-517)
        E += pair_pote(g.c12, g.c6, g.eps, rs);
        if (!(ent.y & 1)) add_stress(fc, rs, S);
      }
    }
    store_bond_sums<MIXED>(d, nb, b, F, T);
    if (THERMO) d.bpote[b] = E;
  }
  if (THERMO) block_stress(d, S);
}

// Load-balanced variant: one wavefront per 64 bonds.  A bond's candidate
// count varies (0 to ~10), and in the per-bond kernel a wave runs as long
// as its busiest lane.  Here the wave's incidences (contiguous in the CSR)
// are dealt out to the lanes in chunks of kChunk: each lane evaluates its
// pairs against the owner's record (decoded once per bond into LDS, the
// owner found from the entry's owner field) and stages the owner's force
// and torque shares in LDS; every lane then adds its own bond's shares in
// incidence order -- the same values summed in the same order as the
// per-bond kernel, so the results are identical bit for bit.  The mixed
// path's force shares are FP32 values and are staged as such.
constexpr int kChunk = 128;

template <bool MIXED, bool THERMO>
__device__ __forceinline__ void pair_bal_body(const Dev &d, const Geo &g, int64_t blk) {
  using FT = typename std::conditional<MIXED, float, double>::type;
  __shared__ BondRec ownr[64];
  __shared__ FT cF[kChunk][3];
  __shared__ double cT[kChunk][3], cE[THERMO ? kChunk : 1];
  __shared__ int32_t qs[65];
  if (d.ctl->abort) return;
  const int64_t nb = g.nb;
  const int lane = threadIdx.x;
  const int64_t b0 = blk * 64;
  const int64_t b = b0 + lane;
  const bool valid = b < nb;
  qs[lane] = d.inc_start[valid ? b : nb];
  if (lane == 0) qs[64] = d.inc_start[b0 + 64 <= nb ? b0 + 64 : nb];
  if (valid) ownr[lane] = load_rec<MIXED>(d.br, g, b);
  double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0}, E = 0.0;
  double S[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (valid && g.nx > 0) {
    const int32_t x0 = d.xinc_start[b], x1 = d.xinc_start[b + 1];
    for (int32_t q = x0; q < x1; ++q) {
      const int32_t e = d.xinc[q];
      const double *o = d.xf + 9 * (int64_t)(e >> 1);
      const int a = e & 1;
      for (int k = 0; k < 3; ++k) F[k] += a == 0 ? o[k] : -o[k];
      for (int k = 0; k < 3; ++k) T[k] += o[3 + 3 * a + k];
    }
  }
  __syncthreads();
  const int32_t Q0 = qs[0], Q1 = qs[64];
  const int32_t my0 = qs[lane], my1 = qs[lane + 1];
  for (int32_t cs = Q0; cs < Q1; cs += kChunk) {
    const int32_t ce = Q1 < cs + kChunk ? Q1 : cs + kChunk;
    for (int32_t q = cs + lane; q < ce; q += 64) {
      const int2 ent = d.inc[q];
      const BondRec oth = load_rec<MIXED>(d.br, g, ent.x);
      const BondRec own = ownr[(ent.y >> 2) & 63];
      double fc[3], sc, rs, cf[3], ct[3];
      pair_slot<MIXED>(d, g, own, q, ent, oth, fc, &sc, &rs);
      const double sg = (ent.y & 1) ? -1.0 : 1.0;
      for (int k = 0; k < 3; ++k) cf[k] = sg * fc[k];
      pair_share<MIXED>(own.u, sc, cf, ct);
      for (int k = 0; k < 3; ++k) {
        cF[q - cs][k] = (FT)cf[k];
        cT[q - cs][k] = ct[k];
      }
      if (THERMO) {
        cE[q - cs] = pair_pote(g.c12, g.c6, g.eps, rs);
        if (!(ent.y & 1)) add_stress(fc, rs, S);
      }
    }
    __syncthreads();
    const int32_t e = my1 < ce ? my1 : ce;
    for (int32_t q = my0 > cs ? my0 : cs; q < e; ++q) {
      for (int k = 0; k < 3; ++k) F[k] += (double)cF[q - cs][k];
      for (int k = 0; k < 3; ++k) T[k] += cT[q - cs][k];
      if (THERMO) E += cE[q - cs];
    }
    __syncthreads();
  }
  if (valid) store_bond_sums<MIXED>(d, nb, b, F, T);
  if (THERMO) {
    if (valid) d.bpote[b] = E;
    block_stress(d, S);
  }
}

template <bool MIXED, bool THERMO>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_pair_bal(Dev d, Geo g, int xcd,
                                                                                           int64_t blk0) {
  pair_bal_body<MIXED, THERMO>(d, g, blk0 + logical_block(xcd));
}

// Ensemble batches: the balanced kernel over every member's bonds (mixed
// precision, no stress sums); `par` 0 on odd steps, 1 on even ones.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_group_pair(
    const Member *__restrict__ ms, GroupGrid gg, int par) {
  int b0, nblk;
  const Member &M = ms[group_member(gg, blockIdx.x, &b0, &nblk)];
  pair_bal_body<true, false>(M.d[par], M.g, blockIdx.x - b0);
}

int64_t pair_blocks(const Geo &g, const Tune &t) {
  const int bs = t.pair_bal ? 64 : t.pair_block;
  const int64_t blocks = (g.nb + bs - 1) / bs;
  return blocks > 0 ? blocks : 1;
}

void launch_pair(const Dev &d, const Geo &g, int precision, const Tune &t, hipStream_t st,
                 bool thermo) {
  launch_pair_range(d, g, precision, t, st, thermo, 0, g.nb);
}

// Bonds [b_lo, b_hi): b_lo a multiple of kPairAlign, and so is b_hi unless
// it is the last bond.
void launch_pair_range(const Dev &d, const Geo &g, int precision, const Tune &t, hipStream_t st,
                       bool thermo, int64_t b_lo, int64_t b_hi) {
  const int bs = t.pair_bal ? 64 : t.pair_block;
  const int64_t blk0 = b_lo / bs;
  const int64_t nblk = (b_hi - b_lo + bs - 1) / bs;
  if (nblk <= 0) return;
  const dim3 grid((unsigned)nblk);
  const bool fp64 = precision == SC_PREC_FP64;
#define SC_PAIR(KERNEL, BS)                                                                    \
  do {                                                                                         \
    if (fp64 && thermo)                                                                       \
      hipLaunchKernelGGL((KERNEL<false, true>), grid, dim3(BS), 0, st, d, g, t.pair_xcd, blk0); \
    else if (fp64)                                                                             \
      hipLaunchKernelGGL((KERNEL<false, false>), grid, dim3(BS), 0, st, d, g, t.pair_xcd, blk0); \
    else if (thermo)                                                                           \
      hipLaunchKernelGGL((KERNEL<true, true>), grid, dim3(BS), 0, st, d, g, t.pair_xcd, blk0);  \
    else                                                                                       \
      hipLaunchKernelGGL((KERNEL<true, false>), grid, dim3(BS), 0, st, d, g, t.pair_xcd, blk0); \
  } while (0)
  if (t.pair_bal)
    SC_PAIR(k_pair_bal, 64);
  else
    SC_PAIR(k_pair_fused, t.pair_block);
#undef SC_PAIR
}

void launch_group_pair(const Member *dm, const Member *hm, int n, bool odd, hipStream_t st) {
  GroupGrid gg;
  int32_t t = 0;
  for (int m = 0; m < n; ++m) {
    gg.off[m] = t;
    t += (int32_t)((hm[m].g.nb + 63) / 64);
  }
  for (int m = n; m <= kMaxGroup; ++m) gg.off[m] = t;
  if (t <= 0) return;
  hipLaunchKernelGGL(k_group_pair, dim3((unsigned)t), dim3(64), 0, st, dm, gg, odd ? 0 : 1);
}

}  // namespace sc
