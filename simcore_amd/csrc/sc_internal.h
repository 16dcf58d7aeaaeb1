// sc_internal.h -- device-side data layout and kernel launchers of the
// MI355X filament BD path.  Internal to libsimcore_amd.so.
//
// HBM layout of one system (all SoA, filament-major):
//   sites  (Ns):  r[3], r_prev[3], f_rand[3], dr0[3], acc        FP64
//   bonds  (Nb):  record {spos[3], u[3], len, -} (64 B), F[3], T[3],
//                 tension                                          FP64
//   cells  (n^3): count, start (n^3+1)                              int32
//   candidates (cap): pairs int2, sticky slot F (double4 or float4),
//                 incidence CSR (2 per candidate)                   int32
// Bonds of filament f occupy [site_off[f]-f, site_off[f+1]-f-1), so
// "shares a site" (Bond::HasNeighbor) is |b1-b2| == 1 within a filament.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/simcore_amd.h"

namespace sc {

// Device-resident control block.  Every kernel of a step reads it, so a
// whole step (rebuild branch included) can be replayed from a hipGraph
// without host round-trips.
struct Ctl {
  int64_t i_step;          // completed steps
  int64_t n_rebuilds;      // completed rebuilds
  int64_t nc;              // live candidates
  int64_t nc_need;         // candidates the last emission scan asked for
  int64_t cur_step;        // step being executed
  unsigned long long dr_bits;  // max per-site accumulator of the last full step
  int32_t abort;           // nonzero: every kernel returns at once
  int32_t err;             // SC_ERR_* when abort was raised by an error
  int32_t rebuild;         // the current step rebuilds the candidate list
  int32_t midstep;         // the current step is a midstep
  int32_t force_rebuild;   // host request (load time)
  int32_t overflow;        // candidate capacity exceeded
  int32_t xl_dirty;        // crosslink bonds changed: rebuild their lists
  int32_t pad_;
  int64_t xl_events;       // crosslink rebinding events so far
};

// Per-system constants passed by value to kernels.
struct Geo {
  double h[9];
  double hinv[9];
  double c12, c6, rcut2, fcut, eps;
  int thermo;              // pair pass also sums stress and bond energies
  double delta;
  double thr;              // 0.25 * cell_len^2
  int n1d;
  int occ_words;           // 32-bit occupancy words per cell column
  int stoch;
  int compact;             // mixed precision: 32-byte bond records (sc_math.h)
  double inv_n1d;
  uint64_t seed;
  // caller-supplied noise (sc_set_noise): xi of midsteps [xi_m0, xi_m0 +
  // xi_nm), 3 per caller site, site-major; null: the counter-based stream
  const double *xi_buf;
  int64_t xi_m0, xi_nm;
  int32_t nf;
  int64_t ns, nb, n_cells, cap;
  // crosslinks
  int64_t nx;
  double xk, xr0, xkc;     // spring, rest length, compression spring
  double xpoff, xrb2;      // unbinding probability per full step, range^2
  uint64_t xseed;
  int xreach;              // cells searched on each side when rebinding
  // order of a bond's 14 home cells (its own cell: code 13; mirrored
  // half-shell cell q: code q) by cell key, 4 bits per entry, first in the
  // low bits; indexed by which of the cell's coordinates lie on a periodic
  // face: (x == 0 ? 0 : x == n-1 ? 2 : 1) * 6 + (y likewise) * 2 + (z > 0)
  uint64_t home_order[18];
};

// Device pointers (a plain struct so it can be passed by value).
// Per-wave maxima of the displacement accumulator go to kDrSlots slots
// (atomicMax, wave index modulo kDrSlots) rather than one word, so that
// thousands of waves do not serialise on one address; the commit folds the
// slots into Ctl::dr_bits and clears them.
constexpr int kDrSlots = 256;

struct Dev {
  Ctl *ctl;
  unsigned long long *dr_slots;  // kDrSlots, zero between steps
  // filaments
  int32_t *site_off;       // nf+1
  int32_t *noise_off;      // nf: site index the noise stream is keyed by
  double *fpar;            // nf*8: bl, kappa, gpar, gperp, sig_par, sig_perp, diam, 0
  int32_t *norm;           // nf*2: normalize_switch_, n_normalize_
  int32_t *bond_fil;       // nb
  // sites
  double *r, *rp, *fr, *dr0, *acc, *acc_prev;
  // bonds
  // bond records, two buffers: a step reads the state's records (br) and
  // its integrator writes the next state's (br_out), so the pair pass and
  // the integrator of different filaments may run at the same time.
  // Outside a step both point at the current state's buffer.
  double *br;              // nb*8: {s.x, s.y, s.z, u.x, u.y, u.z, len, 0}
  double *br_out;
  double *bf, *bt, *tension;  // bf, bt: FP64 precision only
  float *bf32, *bt32;          // per-bond force / torque sums, mixed precision
  // cell list
  // cell list of the last build (CellList::UpdateCellList,
  // src/cell_list.cpp:183-199): bonds sorted by (cell key, bond index) --
  // a counting sort by cell column (x, y) followed by a rank sort of each
  // column in LDS -- with the run of every occupied cell and one
  // occupancy bit per cell
  int32_t *keys;           // nb: cell key of each bond
  int32_t *col_count;      // n1d^2: column populations (zero between builds)
  int32_t *col_start;      // n1d^2 + 1: column starts in sorted
  int32_t *col_rank;       // nb: a bond's place in its column (build scratch)
  int32_t *col_tmp;        // nb: bonds grouped by column, unordered
  int32_t *sorted, *skey;  // nb: bonds in (key, bond) order, their keys
  int2 *cell_run;          // n_cells: [start, end) of an occupied cell
  uint32_t *cell_occ;      // n1d^2 * occ_words: bit z of column (x, y)
  uint16_t *occ_pre;       // n1d^2 * occ_words: occupied cells in the
                           // column's words before this one
  int32_t *col_ncell;      // n1d^2: occupied cells of each column
  int32_t *cell_first;     // nb: per column, at its range of `sorted`, the
                           // first entry of each occupied cell in z order
  // candidates: per bond, its incidences in candidate order (CSR over
  // inc_start), each with its own copy of the pair's sticky force -- both
  // owners compute the same bits, so the copies agree, and a bond's slot
  // traffic follows its list.  The candidate list itself is implicit (the
  // reference's order is recovered from the lists for the taps).
  int32_t *deg, *inc_start;
  int2 *nrun;              // per bond, rebuild scratch: kRunSlab entries (own
                           // cell's run, the first six occupied stencil
                           // cells' runs, the stencil occupancy mask: bits
                           // 0-12 half shell, 13-25 mirrored)
  int2 *inc;               // {partner bond, role | written << 1 | (owner
                           // mod 64) << 2}: role 1 when the bond is obj2;
                           // written once the slot holds a force (a slot
                           // not written since the rebuild is 0)
  double4 *islot64;        // sticky force per incidence {Fx,Fy,Fz,r}, FP64
  float4 *islot32;         // the same in FP32 (mixed precision)
  // scan scratch
  int32_t *scan_tmp;
  // crosslinks: anchors (2 per crosslink), stored forces (9 per
  // crosslink: F, torque on anchor 0's bond, on anchor 1's bond), per-bond
  // lists of {crosslink << 1 | anchor} in crosslink order
  int32_t *xb;
  double *xlam, *xf;
  int32_t *xinc_start, *xinc, *xdeg;
  int32_t *bond_caller;    // storage bond -> caller bond index
  // thermodynamics (sc_thermo_enable): per pair-pass block, the sum of the
  // pair stress tensors xx yy zz xy xz yz over the steps since the last read;
  // per bond, its pair potential energy of the last step
  double *stress_acc, *bpote;
};

// Launch shapes the autotuner chooses among; none changes a result.
struct Tune {
  int pair_block = 256;  // threads per block of the pair kernel
  int pair_xcd = 0;      // 1: consecutive bond ranges stay on one XCD
  int pair_bal = 0;      // 1: incidences dealt evenly over a wave's lanes
  int int_waves = 1;     // waves per block of the register integrator
  int int_layout = 1;    // up to 64 sites: 0 one site per lane, 1 two (the
                         // two round differently; not changed by tuning);
                         // mixed path: 2 packed lane segments, one launch
                         // for every filament (129-256 sites in its first
                         // waves), 3 one launch per lane-group width
  int int_streams = 0;   // length bins spread over this many side streams (0-3)
  int parts = 1;         // >1: the bonds are cut into this many ranges; the
                         // integrator of a range's filaments runs on a side
                         // stream beside the pair pass of the next range
  int early_long = 0;    // 1: the pair pass is cut after the bonds of the
                         // filaments above 128 sites (stored first), whose
                         // integrator then runs beside the rest of the pass
};
constexpr int kMaxParts = 8;
constexpr int kRunSlab = 8;  // cell runs kept per bond between the list passes
constexpr int64_t kPairAlign = 512;  // range cuts: a multiple of every block size

// Length bins of the integrator, as the step loop hands them over.
constexpr int kMaxBins = 6;
struct BinDesc {
  const int32_t *list;  // filament indices of the bin (packed: lane table)
  int count;            // filaments (packed: wavefronts)
  int maxn;             // 8, 16, 32, 64, 128 or 256 sites; 0: packed waves
  int width;            // packed: lane-group width of the waves (16/32/64)
};

// Grouped launches: one launch steps several independent systems (a batch
// of the ensemble runner), each with its own buffers and constants.  Block
// b of a grouped launch belongs to member m with off[m] <= b < off[m + 1];
// the entries past the last member repeat the total.  Each member's blocks
// run the single-system kernel's code with the member's block index and
// block count, so a member's results are those of its own launches.
constexpr int kMaxGroup = 32;
struct Member {
  Dev d[2];                 // [0]: odd steps (read rec[0], write rec[1]); [1]: even
  Geo g;
  const int32_t *pack_tab;  // packed integrator lane table
  int pack_waves, cs1, cs2; // its waves; first wave of width 32, of width 64
  const int32_t *long_list; // filaments of 129-256 sites (one wave each)
  int n_long;
};
struct GroupGrid {
  int32_t off[kMaxGroup + 1];
};

// Member of block `bid`, its first block and block count: a uniform scan
// over the offsets (constant indices, so they stay kernel arguments).
__device__ __forceinline__ int group_member(const GroupGrid &gg, int bid, int *b0, int *nblk) {
  int m = 0, lo = 0, hi = gg.off[1];
#pragma unroll
  for (int i = 1; i < kMaxGroup; ++i) {
    const bool past = bid >= gg.off[i];
    m = past ? i : m;
    lo = past ? gg.off[i] : lo;
    hi = past ? gg.off[i + 1] : hi;
  }
  *b0 = lo;
  *nblk = hi - lo;
  return m;
}

// Grouped launchers (sc_kernels.hip, sc_pair.hip, sc_integrate_long.hip):
// `dm` the device copy of the members, `hm` the host copy (sizes), n
// members, `odd` the parity of the step all of them take.
void launch_group_rebuild(const Member *dm, const Member *hm, int n, hipStream_t st);
void launch_group_pair(const Member *dm, const Member *hm, int n, bool odd, hipStream_t st);
void launch_group_integrate(const Member *dm, const Member *hm, int n, bool odd, hipStream_t st);
void launch_group_commit(const Member *dm, int n, hipStream_t st);
// each member's control block into out[m]
void launch_group_ctl(const Member *dm, int n, Ctl *out, hipStream_t st);


// Per-bond force / torque component q of bond j (summed in FP64 by the pair
// kernel; stored in FP32 on the mixed path).
__device__ __forceinline__ double bond_force(const Dev &d, int64_t nb, int q, int64_t j) {
  return d.bf32 ? (double)d.bf32[q * nb + j] : d.bf[q * nb + j];
}
__device__ __forceinline__ double bond_torque(const Dev &d, int64_t nb, int q, int64_t j) {
  return d.bt32 ? (double)d.bt32[q * nb + j] : d.bt[q * nb + j];
}

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
// grid cap of streaming and scan passes (they loop over their range)
constexpr int kStreamBlocks = 1024;
constexpr int kScanTile = kScanBlock * kScanItems;

enum Guard { kGuardStep = 0, kGuardRebuild = 1, kGuardXlink = 2 };

// Cell-list lookups.  Occupancy bit of cell (x, y, z); run of an occupied
// cell (its members in bond order).
__device__ __forceinline__ bool cell_occupied(const Dev &d, const Geo &g, int x, int y, int z) {
  const uint32_t w = d.cell_occ[((int64_t)x * g.n1d + y) * g.occ_words + (z >> 5)];
  return (w >> (z & 31)) & 1u;
}

// Launchers (sc_kernels.hip).  All are asynchronous on `st`.
void launch_decide(const Dev &d, const Geo &g, hipStream_t st);
void launch_commit(const Dev &d, const Geo &g, hipStream_t st);
void launch_rebuild(const Dev &d, const Geo &g, hipStream_t st);
void launch_cells_from_keys(const Dev &d, const Geo &g, hipStream_t st);
// `thermo`: also sum the pair stress and the bond energies (Geo::thermo)
void launch_pair(const Dev &d, const Geo &g, int precision, const Tune &t,
                 hipStream_t st, bool thermo);
int64_t pair_blocks(const Geo &g, const Tune &t);
void launch_pair_range(const Dev &d, const Geo &g, int precision, const Tune &t, hipStream_t st,
                       bool thermo, int64_t b_lo, int64_t b_hi);  // blocks of one pair launch
// `mid`: the step is a midstep (odd step index)
void launch_integrate(const Dev &d, const Geo &g, const int32_t *bin_list,
                      int bin_count, int bin_maxn, int precision, const Tune &t, bool mid,
                      hipStream_t st);
// register-resident fast-path integrator; false when the bin is not one it
// serves (FP64 precision or more than 64 sites), nothing launched then
bool launch_integrate_reg(const Dev &d, const Geo &g, const int32_t *bin_list,
                          int bin_count, int bin_maxn, int precision, int waves,
                          hipStream_t st);
// several sites per lane (sc_integrate_long.hip): 65..256 sites always,
// up to 64 sites when layout == 1
bool launch_integrate_long(const Dev &d, const Geo &g, const int32_t *bin_list,
                           int bin_count, int bin_maxn, int precision, int waves,
                           int layout, bool mid, hipStream_t st);
// packed layout for the mixed path: filaments of up to 128 sites as lane
// segments of their own width (sc_integrate_long.hip)
void launch_integrate_packed(const Dev &d, const Geo &g, const int32_t *lane_tab, int n_waves,
                             int width, int waves, bool mid, hipStream_t st);
void launch_integrate_packed_all(const Dev &d, const Geo &g, const int32_t *lane_tab,
                                 int n_waves, int cs1, int cs2, int waves, bool mid,
                                 hipStream_t st);
// the same plus the filaments of 129-256 sites (long_list), in one launch
void launch_integrate_all(const Dev &d, const Geo &g, const int32_t *lane_tab, int n_waves,
                          int cs1, int cs2, const int32_t *long_list, int n_long, int waves,
                          bool mid, hipStream_t st);
void launch_bond_init(const Dev &d, const Geo &g, int precision,
                      hipStream_t st);
// crosslinks (sc_crosslink.hip)
void launch_xlink_forces(const Dev &d, const Geo &g, hipStream_t st);
void launch_xlink_events(const Dev &d, const Geo &g, hipStream_t st);
void launch_xlink_lists(const Dev &d, const Geo &g, hipStream_t st);
// analyses (sc_analysis.hip)
void launch_filament_stats(const Dev &d, const Geo &g, double *e2e, double *com,
                           hipStream_t st);
void launch_theta_hist(const Dev &d, const Geo &g, int n_bins, int n_rows,
                       unsigned long long *hist, hipStream_t st);
void launch_zero_i32(int32_t *p, int64_t n, hipStream_t st);
void launch_redo_dr(const Dev &d, const Geo &g, hipStream_t st);
void launch_copy_probe(const void *src, void *dst, int64_t n32, hipStream_t st);
void launch_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *tmp,
                 const Ctl *ctl, int guard, hipStream_t st);
int64_t scan_tmp_elems(int64_t n);

}  // namespace sc
