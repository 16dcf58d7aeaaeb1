/*
 * oracle.h -- CPU restatement of simcore's filament Brownian-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against, and the CPU baseline that bench.py times beside
 * it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (simcore_amd/) never links, loads or
 * includes anything from this directory.
 *
 * Everything here is FP64 and follows the reference's operation order, so
 * that a bit-exact FP64 device path can be compared value-for-value:
 *   cell list      src/cell_list.cpp:9-262, src/cell.cpp:14-70
 *   rebuild rule   src/interaction_manager.cpp:54, 211-231, 284-321,
 *                  src/mesh.cpp:533-556, src/object.cpp:274-287,
 *                  include/simcore/species.hpp:105-120
 *   pair pass      src/interaction_manager.cpp:323-397, 448-517
 *   min distance   src/minimum_distance.cpp:202-344
 *   potentials     include/simcore/wca_potential.hpp:14-48,
 *                  include/simcore/potential_manager.hpp:92-103,
 *                  include/simcore/max_force_potential.hpp:11-40
 *   integrator     src/filament.cpp:394-1048, src/mesh.cpp:168-182, 324-334,
 *                  src/bond.cpp:22-37, src/site.cpp:88-112,
 *                  src/linear_algebra.cpp:3-35, 100-117, 164-197
 *   species loop   src/filament_species.cpp:59-85
 *
 * The reference draws its thermal noise from one GSL generator per filament
 * (src/filament.cpp:527-530).  GSL is absent, so the noise is injected: a
 * counter-based Philox4x32-10 stream keyed by (seed, global site, midstep
 * index) gives xi in (-0.5, 0.5) with 2^-32 resolution, the same variance
 * (1/12) as the reference's uniform_pos()-0.5.
 *
 * Parity pins for each function are listed in the header comment of
 * oracle.c next to the function.
 */
#ifndef SIMCORE_ORACLE_H
#define SIMCORE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  double h[9];              /* unit cell (space_struct.unit_cell)          */
  double h_inv[9];          /* inverse unit cell (space_struct.unit_cell_inv) */
  double system_radius;     /* params.system_radius                        */
  double min_cell_length;   /* CellList::_min_cell_length_                 */
  double delta;             /* params.delta                                */
  double wca_eps, wca_sig;  /* params.wca_eps / wca_sig                    */
  double f_cutoff;          /* params.f_cutoff                             */
  int stoch_flag;           /* params.stoch_flag                           */
  int n_threads;            /* OpenMP threads, <=0: runtime default        */
  uint64_t seed;            /* noise stream key                            */
} orc_params;

typedef struct orc_sys orc_sys;

/* Creates a system from site positions (SoA x[Ns], y[Ns], z[Ns]) of nf
 * filaments laid out contiguously (site_off has nf+1 entries).  Per-filament
 * arrays: nominal bond length, persistence length (bending stiffness with
 * kT=1), parallel/perpendicular friction and diameter.  Performs what
 * Filament::InsertFilament (UpdateBondPositions, UpdatePrevPositions) and the
 * end of Simulation::InsertSpecies (first UpdateInteractions + ZeroDrTot) do.
 * Returns NULL on bad input. */
orc_sys *orc_create(const orc_params *p, int nf, const int32_t *site_off,
                    const double *rx, const double *ry, const double *rz,
                    const double *bond_length, const double *kappa,
                    const double *gamma_par, const double *gamma_perp,
                    const double *diameter);
void orc_destroy(orc_sys *s);

/* Advances n reference steps (i_step += n).  0 on success, -1 if a
 * filament moved more than the reference's unphysical-dr limit, -2 on a
 * self-pair (both are fatal errors in the reference). */
int orc_step(orc_sys *s, int n);

/* ---- state taps (all SoA: x block, y block, z block) ---------------------- */
int64_t orc_n_sites(const orc_sys *s);
int64_t orc_n_bonds(const orc_sys *s);
int64_t orc_n_candidates(const orc_sys *s);       /* live list after cull  */
int64_t orc_rebuild_count(const orc_sys *s);
int64_t orc_step_index(const orc_sys *s);          /* last completed i_step */
int orc_n_cells_1d(const orc_sys *s);
double orc_cell_length(const orc_sys *s);
void orc_get_candidates(const orc_sys *s, int32_t *b1, int32_t *b2);
void orc_get_slot_forces(const orc_sys *s, double *f);  /* 3 x Nc        */
void orc_get_cell_keys(const orc_sys *s, int32_t *keys); /* last rebuild */
void orc_get_sites(const orc_sys *s, double *r, double *r_prev, double *f_rand);
void orc_get_bonds(const orc_sys *s, double *pos, double *spos, double *u,
                   double *len);
void orc_get_bond_forces(const orc_sys *s, double *f, double *t);
void orc_get_tensions(const orc_sys *s, double *tension);   /* per bond  */
void orc_get_dr(const orc_sys *s, double *acc, double *dr_max_per_fil);
void orc_set_noise_buffer(orc_sys *s, const double *xi, int64_t n_mid);
void orc_take_stress(orc_sys *s, double *out);
void orc_get_bond_energies(const orc_sys *s, double *out);
void orc_set_sites(orc_sys *s, const double *r);  /* for forced-error tests */

/* ---- crosslinks (harmonic springs between two bonds) ----------------------- *
 * Force law and application: Crosslink::CalculateTetherForces
 * (src/crosslink.cpp:246-273), MinimumDistance::PointPoint
 * (src/minimum_distance.cpp:71-101), Anchor::UpdateAnchorPositionToBond and
 * ApplyAnchorForces (src/anchor.cpp:113-146, 236-248); schedule of
 * CrosslinkSpecies::UpdatePositions (src/crosslink_species.cpp:218-238):
 * even steps recompute the forces from the current bonds and apply them,
 * odd steps re-apply the stored forces, both before the pair pass and in
 * crosslink order.  Anchors do not move along their bonds.  Binding
 * kinetics are this library's model (the reference's KMC is not
 * available): on even step k each crosslink, with probability
 * p_off = 1 - exp(-k_off dt), releases one anchor (chosen by a draw) and
 * rebinds it to the in-range bond (point-segment distance <= r_bind from
 * the staying anchor, staying bond excluded) with the smallest Philox score
 * at a uniform fraction of its nominal length; its stored force is zeroed.
 * Draws: Philox4x32-10, key = seed, counter = (crosslink, k/2, 0, 1) for the
 * event and (crosslink, k/2, bond, 2) for the scores. */
typedef struct orc_xlink_params {
  double k_spring, rest_length, k_compress; /* k_compress < 0: none      */
  double k_off;                             /* unbinding rate per time   */
  double r_bind;                            /* rebinding range           */
  uint64_t seed;
} orc_xlink_params;
/* bonds: 2*nx (anchor 0, anchor 1 per crosslink), lambda: 2*nx distances
 * from the bond's first site.  Stored forces start at zero. */
int orc_load_crosslinks(orc_sys *s, int64_t nx, const int32_t *bonds,
                        const double *lambda, const orc_xlink_params *p);
/* bonds/lambda as above; force: 9*nx (F, torque on anchor 0's bond, torque
 * on anchor 1's bond) stored from the last even step; any may be NULL. */
void orc_get_crosslinks(const orc_sys *s, int32_t *bonds, double *lambda,
                        double *force);
int64_t orc_xlink_events(const orc_sys *s);
/* force of one spring between two anchor points (for pinning tests):
 * out = F[3] on anchor 0; returns the spring length. */
double orc_xlink_force(const double *h, const double *h_inv, const double *p0,
                       const double *p1, double k, double r0, double kc,
                       double *out);

/* ---- unit-level restatements (for pinning tests) --------------------------- */
/* MinimumDistance::Sphero.  out: r_min[3], r_min_mag2, contact1[3],
 * contact2[3] (10 doubles). */
void orc_min_distance(const double *h, const double *s1, const double *u1,
                      double l1, const double *s2, const double *u2, double l2,
                      double *out);
/* Potential for one pair given dr[3], dr_mag2, contact1[3], contact2[3].
 * Writes force[3]; returns 1 if the force was written, 0 if the pair is
 * beyond the cutoff or the max-force fallback aborted (sticky in both). */
int orc_pair_force(double c12, double c6, double rcut2, double f_cutoff,
                   const double *dr, double dr_mag2, const double *c1,
                   const double *c2, double *force);
/* tridiagonal_solver: in-place Thomas on n unknowns (c is modified). */
void orc_tridiag(int n, const double *a, const double *b, double *c, double *d);
/* 13-cell half-shell stencil of cell (x,y,z), reference insertion order. */
int orc_cell_neighbors(int n1d, int x, int y, int z, int32_t *out_keys);
/* Cell coordinates of a scaled position, packed as (x*n+y)*n+z. */
int32_t orc_cell_key(int n1d, const double *spos);
/* Philox4x32-10 block. */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]);
/* xi triple of global site `site` at midstep index `m` */
void orc_noise(uint64_t seed, int64_t site, int64_t m, double xi[3]);
/* Filament::CalculateBendingForces for one filament: u is the bond
 * orientation array (3 x nb SoA, nb = n_sites-1); writes site forces
 * (3 x n_sites SoA, overwritten, not accumulated) and g_inv / k_eff. */
void orc_bending_forces(int n_sites, const double *u, double kappa,
                        double bond_length, double *f_site, double *k_eff);
/* Filament::ApplyInteractionForces for one filament: per-bond orientation,
 * force and torque (3 x nb SoA each, nb = n_sites-1) to site forces
 * (3 x n_sites SoA, overwritten). */
void orc_bond_site_forces(int n_sites, double bond_length, const double *u, const double *f,
                          const double *t, double *f_site);

#ifdef __cplusplus
}
#endif
#endif
