// sc_analysis.hip -- on-device filament analyses: end-to-end vectors,
// centres of mass (FilamentSpecies::RunMse2eAnalysis,
// src/filament_species.cpp:569-593) and the cos(theta) histogram of
// consecutive bonds (RunThetaAnalysis, :595-610).  Read-only over the state;
// cheap next to a step, so long statistical runs can sample often.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_math.h"

namespace sc {

namespace {

__global__ __launch_bounds__(256) void k_filament_stats(Dev d, Geo g, double *e2e,
                                                        double *com) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nf) return;
  const int64_t s0 = d.site_off[f], s1 = d.site_off[f + 1];
  const int64_t ns = g.ns, nf = g.nf;
  for (int q = 0; q < 3; ++q) {
    const double *r = d.r + q * ns;
    if (e2e) e2e[q * nf + f] = r[s1 - 1] - r[s0];
    if (com) {
      double s = 0.0;
      for (int64_t i = s0; i < s1; ++i) s += r[i];
      com[q * nf + f] = s / (double)(s1 - s0);
    }
  }
}

// one thread per bond j with a successor in its filament: angle index
// (j - first bond of the filament), cos = u_j . u_{j+1}
__global__ __launch_bounds__(256) void k_theta_hist(Dev d, Geo g, int n_bins, int n_rows,
                                                    unsigned long long *hist) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.nb) return;
  const int f = d.bond_fil[j];
  const int64_t b0 = d.site_off[f] - f;
  const int64_t b1 = d.site_off[f + 1] - f - 1;  // one past the last bond
  if (j + 1 >= b1) return;
  const int row = (int)(j - b0);
  if (row >= n_rows) return;
  double s[3], a[3], b[3], l;
  rec_read(d.br, g, j, s, a, &l);
  rec_read(d.br, g, j + 1, s, b, &l);
  const double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  int bin = (int)floor((1.0 + c) * (n_bins / 2));
  if (bin >= n_bins) bin = n_bins - 1;
  if (bin < 0) bin = 0;
  atomicAdd(&hist[(int64_t)row * n_bins + bin], 1ull);
}

}  // namespace

void launch_filament_stats(const Dev &d, const Geo &g, double *e2e, double *com,
                           hipStream_t st) {
  const int blocks = (g.nf + 255) / 256;
  hipLaunchKernelGGL(k_filament_stats, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, d, g,
                     e2e, com);
}

void launch_theta_hist(const Dev &d, const Geo &g, int n_bins, int n_rows,
                       unsigned long long *hist, hipStream_t st) {
  const int64_t blocks = (g.nb + 255) / 256;
  hipLaunchKernelGGL(k_theta_hist, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(256), 0, st,
                     d, g, n_bins, n_rows, hist);
}

}  // namespace sc
