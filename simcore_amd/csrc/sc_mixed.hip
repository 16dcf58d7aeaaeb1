// sc_mixed.hip -- mixed-precision pair pass.
//
// The periodic minimum image is formed in FP64 from the FP64 scaled
// positions (the box is up to ~250 sigma, so FP32 scaled coordinates would
// cost ~1e-5 sigma in dr), then rounded to FP32 for the segment-segment
// distance and the WCA force.  Candidates whose orientations are nearly
// parallel (1-(u1.u2)^2 < kParallelSwitch, where the FP32 line parameters
// lose digits) or that overlap to dr^2 < 1e-6 are compacted with a
// wavefront ballot into an LDS queue and redone in FP64 with the full-
// precision reference algorithm, so the queue runs with all lanes busy
// instead of serialising every wave on a divergent branch.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

constexpr float kParallelSwitch = 2.0e-2f;

__device__ __forceinline__ float fabs_(float x) { return x < 0.f ? -x : x; }
__device__ __forceinline__ float fsign_(float a, float b) {
  return b >= 0.f ? fabsf(a) : -fabsf(a);
}

// FP32 MinimumDistance::Sphero (src/minimum_distance.cpp:231-343) given the
// FP32 minimum-image separation.  Returns false when the FP64 path must be
// used instead.
__device__ __forceinline__ bool sphero32(const float *dr, const float *u1,
                                         float hl1, const float *u2, float hl2,
                                         float *rmin, float *rm2, float *lam_o,
                                         float *mu_o) {
  const float a = dr[0] * u1[0] + dr[1] * u1[1] + dr[2] * u1[2];
  const float b = dr[0] * u2[0] + dr[1] * u2[1] + dr[2] * u2[2];
  const float c = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
  const float denom = 1.0f - c * c;
  if (denom < kParallelSwitch) return false;
  float lambda = (a - c * b) / denom;
  float mu = (-b + c * a) / denom;
  const float lm = fabs_(lambda), mm = fabs_(mu);
  if (lm > hl1 && mm > hl2) {
    float la = fsign_(hl1, lambda);
    float ma = -b + la * c;
    if (fabs_(ma) > hl2) ma = fsign_(hl2, ma);
    float mb = fsign_(hl2, mu);
    float lb = a + mb * c;
    if (fabs_(lb) > hl1) lb = fsign_(hl1, lb);
    float m2a = 0.f, m2b = 0.f, ra[3], rb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ra[i] = dr[i] - la * u1[i] + ma * u2[i];
      rb[i] = dr[i] - lb * u1[i] + mb * u2[i];
      m2a += ra[i] * ra[i];
      m2b += rb[i] * rb[i];
    }
    if (m2a < m2b) {
      lambda = la;
      mu = ma;
    } else {
      lambda = lb;
      mu = mb;
    }
  } else if (lm > hl1) {
    lambda = fsign_(hl1, lambda);
    mu = -b + lambda * c;
    if (fabs_(mu) > hl2) mu = fsign_(hl2, mu);
  } else if (mm > hl2) {
    mu = fsign_(hl2, mu);
    lambda = a + mu * c;
    if (fabs_(lambda) > hl1) lambda = fsign_(hl1, lambda);
  }
  float m2 = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rmin[i] = dr[i] - lambda * u1[i] + mu * u2[i];
    m2 += rmin[i] * rmin[i];
  }
  if (m2 < 1.0e-6f) return false;
  *rm2 = m2;
  *lam_o = lambda;
  *mu_o = mu;
  return true;
}

// One candidate fully in FP64 (same code as the FP64 path).
__device__ __forceinline__ void pair_fp64_one(const Dev &d, const Geo &g,
                                              int64_t c) {
  const int64_t nb = g.nb;
  const int2 pr = d.pairs[c];
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
      d.slot32[c] = make_float4((float)f[0], (float)f[1], (float)f[2], 0.f);
  }
  d.lm32[c] = make_float2((float)lam, (float)mu);
}

__global__ __launch_bounds__(256) void k_pair32(Dev d, Geo g) {
  if (d.ctl->abort) return;
  __shared__ int32_t queue[256];
  __shared__ int qn;
  const int64_t nc = d.ctl->nc;
  const int64_t nb = g.nb;
  const float rc2 = (float)g.rcut2;
  const float c12 = (float)g.c12, c6 = (float)g.c6, fcut = (float)g.fcut;
  const int64_t tile = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nc; base += tile) {
    if (threadIdx.x == 0) qn = 0;
    __syncthreads();
    const int64_t c = base + threadIdx.x;
    bool fallback = false;
    if (c < nc) {
      const int2 pr = d.pairs[c];
      const int32_t b1 = pr.x, b2 = pr.y;
      float dr[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double ds = d.spos[k * nb + b2] - d.spos[k * nb + b1];
        ds -= (double)nint(ds);
        dr[k] = (float)(g.h[4 * k] * ds);
      }
      const float4 q1 = d.ub32[b1], q2 = d.ub32[b2];
      const float u1[3] = {q1.x, q1.y, q1.z}, u2[3] = {q2.x, q2.y, q2.z};
      float rmin[3], rm2, lam, mu;
      if (sphero32(dr, u1, q1.w, u2, q2.w, rmin, &rm2, &lam, &mu)) {
        if (!(rm2 > rc2)) {
          const float rinv = 1.0f / sqrtf(rm2);
          const float rinv2 = rinv * rinv;
          const float r6 = rinv2 * rinv2 * rinv2;
          float ffac = -(12.0f * c12 * r6 - 6.0f * c6) * r6 * rinv;
          if (fabs_(ffac) > fcut) ffac = (ffac >= 0.f ? 1.f : -1.f) * fcut;
          d.slot32[c] = make_float4(ffac * rmin[0] * rinv, ffac * rmin[1] * rinv,
                                    ffac * rmin[2] * rinv, 0.f);
        }
        d.lm32[c] = make_float2(lam, mu);
      } else {
        fallback = true;
      }
    }
    // wavefront ballot compaction of the FP64 fallbacks into the LDS queue
    const unsigned long long m = __ballot(fallback);
    const int lane = threadIdx.x & 63;
    int wbase = 0;
    if (lane == 0 && m) wbase = atomicAdd(&qn, __popcll(m));
    wbase = __shfl(wbase, 0, 64);
    if (fallback) {
      const int rank = __popcll(m & ((1ull << lane) - 1ull));
      queue[wbase + rank] = (int32_t)threadIdx.x;
    }
    __syncthreads();
    const int nq = qn;
    if (threadIdx.x < nq) pair_fp64_one(d, g, base + queue[threadIdx.x]);
    __syncthreads();
  }
}

// ApplyPairInteractions as a per-bond segmented sum in candidate order,
// FP32 slots accumulated in FP64.
__global__ __launch_bounds__(256) void k_reduce32(Dev d, Geo g) {
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
    const float4 fs = d.slot32[c];
    const double fc[3] = {fs.x, fs.y, fs.z};
    const float2 lm = d.lm32[c];
    const double sc = (ent & 1) ? (double)lm.y : (double)lm.x;
    double con[3], t[3];
    for (int k = 0; k < 3; ++k) con[k] = sc * ub[k];
    cross3(con, fc, t);
    if ((ent & 1) == 0) {
      for (int k = 0; k < 3; ++k) F[k] += fc[k];
      for (int k = 0; k < 3; ++k) T[k] += t[k];
    } else {
      for (int k = 0; k < 3; ++k) F[k] -= fc[k];
      for (int k = 0; k < 3; ++k) T[k] -= t[k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    d.bf[k * nb + b] = F[k];
    d.bt[k * nb + b] = T[k];
  }
}

void launch_pair32(const Dev &d, const Geo &g, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_pair32, dim3(grid), dim3(256), 0, st, d, g);
}

void launch_reduce32(const Dev &d, const Geo &g, hipStream_t st) {
  int64_t b = (g.nb + 255) / 256;
  hipLaunchKernelGGL(k_reduce32, dim3((unsigned)(b > 0 ? b : 1)), dim3(256), 0,
                     st, d, g);
}

}  // namespace sc
