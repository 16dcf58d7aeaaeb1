/*
 * simcore_amd.h -- C ABI of the MI355X (gfx950) filament Brownian-dynamics
 * hot path.  Plain pointers and sizes only; every call returns an int status
 * (SC_OK = 0, negative = error class) and never throws.  sc_last_error()
 * holds the message of the last failure on a handle.
 *
 * A handle owns one system's device state on one HIP stream.  Caller arrays
 * are borrowed for the duration of a call (copied in or out).  One host
 * thread per handle at a time.
 *
 * Site and bond arrays are SoA: x block, y block, z block (3*N doubles).
 * Sites of filament f are [site_off[f], site_off[f+1]); its bonds are
 * [site_off[f]-f, site_off[f+1]-f-1).
 */
#ifndef SIMCORE_AMD_H
#define SIMCORE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SC_ABI_VERSION 1

/* status codes */
#define SC_OK 0
#define SC_ERR_ARG (-1)        /* bad argument / unsupported configuration   */
#define SC_ERR_HIP (-2)        /* HIP runtime failure                        */
#define SC_ERR_STATE (-3)      /* call out of order (e.g. step before load)  */
#define SC_ERR_UNPHYSICAL (-4) /* dr > 10000: the reference throws
                                  runtime_error("Unphysical result")
                                  (include/simcore/species.hpp:114-118)     */
#define SC_ERR_NOMEM (-5)      /* device memory / admission failure          */
#define SC_ERR_PARSE (-6)      /* parameter file could not be parsed         */

/* arithmetic of the pair path */
#define SC_PREC_FP64 0  /* everything FP64, bit-compatible with the CPU path */
#define SC_PREC_MIXED 1 /* FP64 pair geometry with the hardware FP32
                           reciprocal / reciprocal square root, 32-byte bond
                           records (32-bit fixed-point offset inside the
                           bond's cell, octahedral unit vector, two-float
                           length), FP32 sticky forces, per-bond sums and
                           random forces, FP64 tensions and positions; at
                           most 1024 cells per side                          */

typedef struct sc_system *sc_handle;

/* System parameters.  Replaces the system_parameters fields the hot path
 * reads (include/simcore/parameters.hpp:157-235) plus the derived cell-list
 * and unit-cell quantities (src/cell_list.cpp:9-64, src/space.cpp:85-138). */
typedef struct sc_params {
  int n_dim;               /* must be 3                                    */
  int n_periodic;          /* must be 3 (fully periodic box)               */
  double unit_cell[9];     /* space_struct.unit_cell (orthorhombic)        */
  double unit_cell_inv[9]; /* space_struct.unit_cell_inv                   */
  double system_radius;    /* params.system_radius                         */
  double min_cell_length;  /* CellList::_min_cell_length_ (max over the
                              SetMinCellLength calls)                       */
  double delta;            /* params.delta (one Delta t = two steps)       */
  double wca_eps;          /* params.wca_eps                               */
  double wca_sig;          /* params.wca_sig                               */
  double f_cutoff;         /* params.f_cutoff                              */
  int stoch_flag;          /* params.stoch_flag                            */
  int precision;           /* SC_PREC_FP64 or SC_PREC_MIXED                */
  int use_graph;           /* capture the step in a hipGraph               */
  int spatial_order;       /* 1: store filaments in space-filling-curve order
                              of their centres (load time) so that pair
                              partners sit close in memory; getters still
                              use the caller's order and the noise stream is
                              keyed by the caller's site indices, but the
                              candidate list ORDER (hence FP64 summation
                              order) differs from the reference's          */
  uint64_t seed;           /* thermal-noise stream key                     */
  double capacity_factor;  /* initial candidate capacity / bonds (0: 4);
                              < 0: an absolute initial capacity of
                              -capacity_factor candidates.  The capacity
                              grows on overflow (the step is re-run, results
                              unchanged), up to 2^30 - 1 candidates        */
} sc_params;

/* Per-filament parameters (Filament::SetParameters / SetDiffusion,
 * src/filament.cpp:6-84, 320-348). */
typedef struct sc_filament_params {
  double bond_length;      /* Filament::bond_length_                       */
  double persistence_length; /* bending_stiffness_ (kT = 1)               */
  double gamma_par;        /* friction_par_                                */
  double gamma_perp;       /* friction_perp_                               */
  double diameter;         /* diameter_                                    */
} sc_filament_params;

int sc_abi_version(void);

/* Replaces Simulation::InitSimulation's space / interaction-manager /
 * cell-list set-up: Space::Init (src/space.cpp:15-83),
 * InteractionManager::Init + InitInteractions
 * (src/interaction_manager.cpp:4-56), CellList::Init (src/cell_list.cpp:
 * 17-64), WCAPotential::Init (include/simcore/wca_potential.hpp:34-48). */
int sc_create(const sc_params *params, sc_handle *out);

/* Replaces the state FilamentSpecies holds after Simulation::InsertSpecies
 * (src/simulation.cpp:261-405): Filament::InsertFilament's
 * UpdateBondPositions + UpdatePrevPositions (src/filament.cpp:225-249) and
 * the final ResetCellList + CheckUpdateObjects -> UpdateInteractions +
 * ZeroDrTot (src/simulation.cpp:400-403).  r is SoA (3*Ns doubles). */
int sc_load_filaments(sc_handle h, int32_t n_filaments,
                      const int32_t *site_off, const double *r,
                      const sc_filament_params *fil);

/* Replaces n iterations of Simulation::RunSimulation's step body
 * (src/simulation.cpp:37-62): ZeroForces, InteractionManager::Interact
 * (src/interaction_manager.cpp:63-95) and FilamentSpecies::UpdatePositions
 * (src/filament_species.cpp:59-111), honouring the midstep parity. */
int sc_step(sc_handle h, int32_t n_steps);
/* sc_step split in two, for callers that overlap host work or other systems
 * with the device (the reference's loop is synchronous; the split itself
 * has no counterpart there).  sc_step_async enqueues n steps behind any
 * already enqueued and returns at once; sc_step_poll never blocks: *done is
 * 1 when every enqueued step has completed.  Candidate-capacity growth and
 * the "Unphysical result" error are handled and reported by sc_step_poll.
 * The getters wait for enqueued work; sc_set_sites refuses while steps are
 * in flight. */
int sc_step_async(sc_handle h, int32_t n_steps);
int sc_step_poll(sc_handle h, int *done);

/* Object::GetPosition / GetPrevPosition of sites (include/simcore/
 * object.hpp:82-88); f_rand is Site::random_force_ (src/site.cpp:103-107),
 * FP32-rounded on the mixed path.  Any pointer may be NULL. */
int sc_get_sites(sc_handle h, double *r, double *r_prev, double *f_rand);
/* Bond state after Bond::ReInit (src/bond.cpp:22-37): centre, scaled
 * position, unit orientation, length.  Any pointer may be NULL.  Mixed
 * precision: the scaled position, orientation and length as the compact
 * record holds them (position within 2^-32 of a cell, orientation within
 * about 1e-9, length to about 1e-14 relative); the centre is computed from
 * the FP64 sites. */
int sc_get_bonds(sc_handle h, double *pos, double *spos, double *u,
                 double *len);
/* Per-bond force/torque after ApplyPairInteractions of the last step
 * (src/interaction_manager.cpp:500-517). */
int sc_get_bond_forces(sc_handle h, double *f, double *t);
/* Filament::tensions_ of the last CalculateTensions, one per bond
 * (src/filament.cpp:778-838). */
int sc_get_tensions(sc_handle h, double *tension);
/* CellList::FindCellCoords of every bond at the last rebuild, packed
 * (x*n + y)*n + z (src/cell_list.cpp:139-157). */
int sc_get_cell_keys(sc_handle h, int32_t *keys);
/* InteractionManager::pair_interactions_ (live list, reference order):
 * obj1/obj2 bond indices.  *n receives the count; arrays may be NULL. */
int sc_get_candidates(sc_handle h, int32_t *b1, int32_t *b2, size_t cap,
                      size_t *n);
/* Interaction::force of each live candidate (sticky slot), 3*n doubles. */
int sc_get_slot_forces(sc_handle h, double *f, size_t cap);
/* Displacement bookkeeping for the rebuild rule (src/mesh.cpp:533-556,
 * src/object.cpp:274-287): per-site accumulator. */
int sc_get_dr_state(sc_handle h, double *acc);
/* Counters: steps done, rebuilds done, live candidates, cells per side. */
int sc_get_counters(sc_handle h, int64_t *i_step, int64_t *n_rebuilds,
                    int64_t *n_candidates, int32_t *n_cells_1d);
/* Overwrite site positions (host-side perturbation, e.g. to exercise the
 * unphysical-dr error path).  Bond state is re-initialised from them. */
int sc_set_sites(sc_handle h, const double *r);

/* Thermodynamics (the thermo_flag output of the reference).  While enabled,
 * every pair pass also sums each live candidate's stress -dr_i F_j
 * (WCAPotential / MaxForcePotential::CalcPotential, include/simcore/
 * wca_potential.hpp:28-31, include/simcore/max_force_potential.hpp:35-38;
 * sticky like the force) into InteractionManager::stress_ and each bond's
 * pair potential energy (Object::AddPotential of both bonds,
 * src/interaction_manager.cpp:500-531).  sc_thermo_enable starts a fresh
 * stress sum.  sc_calculate_pressure is InteractionManager::
 * CalculatePressure (src/interaction_manager.cpp:534-559): pressure tensor
 * delta_ij n_bonds / V + stress_ij / (V n_thermo), pressure its trace / 3,
 * V the unit cell's volume; it resets the sum.  sc_get_bond_energies: each
 * bond's energy of the last step.  Both return SC_ERR_STATE while disabled.
 * Mixed precision sums the stress from FP32 slot values (about 1e-7
 * relative). */
int sc_thermo_enable(sc_handle h, int on);
int sc_calculate_pressure(sc_handle h, int32_t n_thermo, double *pressure_tensor,
                          double *pressure, double *volume);
int sc_get_bond_energies(sc_handle h, double *pote);

/* Thermal noise.  The reference draws xi = RandomUniform() - 0.5 three
 * times per site on every midstep from each filament's GSL generator
 * (Filament::ConstructUnprojectedRandomForces, src/filament.cpp:518-548).
 * SC_NOISE_COUNTER (the default): xi from a counter-based stream keyed by
 * (seed, caller site index, midstep index); `seed` replaces the key, xi and
 * n_xi are ignored.  SC_NOISE_BUFFER: the caller's pre-drawn values for the
 * next n_xi / (3 * n_sites) midsteps, 3 per site (x, y, z), sites in the
 * caller's order, midstep-major; stepping into a midstep the buffer does
 * not hold returns SC_ERR_STATE.  The buffer is copied; it is not part of a
 * checkpoint, and the autotuner's trials draw from the counter stream. */
#define SC_NOISE_COUNTER 0
#define SC_NOISE_BUFFER 1
int sc_set_noise(sc_handle h, int mode, uint64_t seed, const double *xi, size_t n_xi);

/* On-device analyses of the current state.  sc_filament_stats: per filament
 * the end-to-end vector head - tail (FilamentSpecies::RunMse2eAnalysis,
 * src/filament_species.cpp:569-593) and the centre of mass, 3*nf doubles
 * each (SoA); either pointer may be NULL.  sc_theta_histogram: cos(theta)
 * of consecutive bonds binned as in RunThetaAnalysis (:595-610),
 * bin = floor((1 + cos) * (n_bins / 2)) clamped to [0, n_bins-1], one row
 * of n_bins per angle index i (bonds i, i+1 of a filament), rows
 * [0, n_rows); counts are ADDED to hist (n_rows * n_bins int64), so calls
 * accumulate samples. */
int sc_filament_stats(sc_handle h, double *e2e, double *com);
int sc_theta_histogram(sc_handle h, int32_t n_bins, int32_t n_rows,
                       int64_t *hist);
/* No counterpart in the reference: sc_theta_histogram for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  It
 * calls no hipMalloc/hipFree: its counters live in a per-GPU device array of
 * at most 65536 counters (n_bins * n_rows; larger shapes return SC_ERR_ARG),
 * cleared at each call, and the counts are ADDED to hist as in
 * sc_theta_histogram.  Every call on one GPU shares that array, so calls on
 * one GPU that overlap in time (two host threads, or handles on different
 * streams without ordering) race. */
int sc_theta_histogram_noalloc(sc_handle h, int32_t n_bins, int32_t n_rows,
                               int64_t *hist);

/* Checkpoint / resume.  The reference's checkpoint (Species::WriteCheckpoints,
 * include/simcore/species.hpp:238-258; Mesh::WriteCheckpoint,
 * src/mesh.cpp:506-514) stores object specs plus GSL RNG states; the noise
 * here is a counter-based stream keyed by the step index, so the complete
 * dynamic state is: sites, previous positions, random forces, displacement
 * bookkeeping, bond state and forces, tensions, normalisation switches,
 * the candidate list with its sticky forces and incidence lists, and the
 * step/rebuild counters.  sc_checkpoint_save with buf == NULL reports the
 * size.  sc_checkpoint_load restores into a handle created with the same
 * parameters and loaded with the same filament layout (checked); stepping
 * then continues bit-identically to the saved handle. */
int sc_checkpoint_save(sc_handle h, void *buf, size_t cap, size_t *size);
int sc_checkpoint_load(sc_handle h, const void *buf, size_t size);

/* Launch shapes (no counterpart in the reference).  None of them changes
 * any result except integrate_layout, which changes the rounding of the
 * mixed-precision integrator.
 *   pair_block          pair kernel threads per block, 64/128/256/512; 0
 *                       selects the load-balanced variant (one wave per 64
 *                       bonds, incidences dealt evenly over its lanes)
 *   pair_xcd            1: consecutive bond ranges on one XCD
 *   integrate_waves     integrator waves per block, 1/2/4
 *   integrate_layout    filaments of up to 64 sites: one (0) or two (1)
 *                       sites per lane in lane groups of a fixed width per
 *                       length bin; 2 (mixed-precision default): two sites
 *                       per lane, every filament of up to 128 sites in a
 *                       lane segment of its own width, packed into shared
 *                       wavefronts, one launch that also carries the
 *                       filaments of 129-256 sites in its first waves
 *                       (same results as 1); 3: the packed filaments in
 *                       one launch per lane-group width (16/32/64), the
 *                       longer ones in a launch of their own
 *   integrate_streams   0-3: the integrator's length bins are spread over the
 *                       system's stream and this many side streams (forked
 *                       and joined inside the step) and run concurrently
 *   pair_parts          0/1: the pair pass, then the integrator; 2-8: the
 *                       bonds are cut into this many ranges, and the
 *                       integrator of the filaments below a cut runs on a
 *                       side stream beside the pair pass of the next range
 *                       (systems without crosslinks; results unchanged)
 *   early_long          1: the pair pass is cut after the bonds of the
 *                       filaments above 128 sites (stored first in spatial
 *                       order), and their integrator runs on a side stream
 *                       beside the rest of the pass (systems without
 *                       crosslinks; results unchanged) */
typedef struct sc_launch_shape {
  int32_t pair_block, pair_xcd, integrate_waves, integrate_layout, integrate_streams;
  int32_t pair_parts, early_long;
} sc_launch_shape;
int sc_set_launch_shape(sc_handle h, const sc_launch_shape *shape);
int sc_get_launch_shape(sc_handle h, sc_launch_shape *shape);

/* Autotuner (no counterpart in the reference): times the launch shapes
 * that leave every result unchanged -- first the pair kernel's (block size
 * or balanced variant, XCD order), then the integrator's (waves per block,
 * side streams, and on the packed layout one launch or one per lane-group
 * width; the sites per lane are kept), then the overlap of the two over
 * bond ranges, and the early integration of the long filaments -- over
 * `steps` steps each from a checkpoint of the current state; the four
 * fastest are timed twice more, each scored by the fastest of its three
 * timings; keeps the best and restores the state, so the trajectory is the
 * one without tuning.  The cell size
 * is not tuned: in the reference it sets the rebuild threshold and the
 * candidate set (src/cell_list.cpp:9-64), i.e. it is part of the model. */
typedef struct sc_tune_report {
  sc_launch_shape shape;  /* the one kept */
  int32_t n_trials;
  double step_ms_start, step_ms_best;
} sc_tune_report;
int sc_autotune(sc_handle h, int32_t steps, sc_tune_report *report);

/* Device bytes this handle holds. */
int sc_device_bytes(sc_handle h, size_t *bytes);
/* Average per-launch duration (ms) of each kernel family recorded with HIP
 * events on the handle's stream while profiling was enabled; families:
 * 0 pair pass with the fused per-bond force scatter (and crosslinks),
 * 1 commit (step counter, next-step decision),
 * 2 integrate (all length bins), 3 rebuild stage (odd steps only). */
int sc_profile_enable(sc_handle h, int on);
int sc_profile_read(sc_handle h, double *avg_ms, int64_t *launches,
                    int n_families);

/* Measurement probe with no counterpart in the reference: launches a plain
 * streaming copy of `bytes` (read once, written once) `repeats` times and
 * reports the average duration; used to calibrate hardware byte counters
 * and as the achievable-bandwidth reference next to the roofline. */
int sc_copy_probe(size_t bytes, int repeats, double *avg_ms);
/* Measurement probe with no counterpart in the reference: runs the pair
 * pass `launches` times back to back on the handle's stream between two
 * HIP events and reports the average duration of one launch (the roofline
 * figure).  The pass only rewrites the same slot values and per-bond sums on
 * an unchanged state, so the system is left as it was. */
int sc_time_pair(sc_handle h, int32_t launches, double *avg_ms);

const char *sc_last_error(sc_handle h);
int sc_destroy(sc_handle h);

/* ---- crosslinks -------------------------------------------------------------- */
/* Harmonic springs between points on two bonds.  Force law, anchor geometry
 * and application: Crosslink::CalculateTetherForces (src/crosslink.cpp:
 * 246-273), MinimumDistance::PointPoint (src/minimum_distance.cpp:71-101),
 * Anchor::UpdateAnchorPositionToBond / ApplyAnchorForces (src/anchor.cpp:
 * 132-146, 236-248).  Schedule of CrosslinkSpecies::UpdatePositions
 * (src/crosslink_species.cpp:218-238): even steps recompute the forces from
 * the current bonds, odd steps re-apply the stored ones; both are added to
 * the bonds before the pair forces, in crosslink order (anchor 0 takes +F,
 * anchor 1 -F).  Anchors do not move along their bonds.  Binding kinetics
 * are this library's model, not the reference's KMC (absent): on even
 * step k each crosslink releases, with probability 1 - exp(-k_off dt), the
 * anchor a draw picks and rebinds it, at a uniform fraction of the nominal
 * bond length, to the bond within r_bind (point-segment distance) of the
 * other anchor that has the smallest Philox score; its stored force is
 * zeroed.  Counter-based draws keep this independent of storage order. */
typedef struct sc_crosslink_params {
  double k_spring;     /* spring constant (kT / sigma^2)                  */
  double rest_length;  /* sigma                                           */
  double k_compress;   /* constant when compressed, < 0: k_spring         */
  double k_off;        /* unbinding rate per unit time (0: static)        */
  double r_bind;       /* rebinding range (sigma)                         */
  uint64_t seed;       /* key of the binding draws                        */
} sc_crosslink_params;

/* bonds: 2*n bond indices (anchor 0, anchor 1 per crosslink, caller bond
 * numbering), lambda: 2*n distances from each bond's first site.  Call
 * after sc_load_filaments and before the first step. */
int sc_load_crosslinks(sc_handle h, int64_t n, const int32_t *bonds,
                       const double *lambda, const sc_crosslink_params *p);
/* Current bonds / lambdas (2*n each) and stored forces (9*n: F, torque on
 * anchor 0's bond, torque on anchor 1's bond); any pointer may be NULL.
 * *events receives the number of rebinding events so far. */
int sc_get_crosslinks(sc_handle h, int32_t *bonds, double *lambda, double *force,
                      int64_t *events);
/* Host generator: n crosslinks between a uniformly chosen bond (uniform
 * fraction along it) and a uniformly chosen other bond within r_bind of
 * that anchor point (uniform fraction). */
int sc_generate_crosslinks(const sc_params *params, int32_t n_filaments,
                           const int32_t *site_off, const double *r,
                           const sc_filament_params *fil, int64_t n,
                           double r_bind, uint64_t seed, int32_t *bonds,
                           double *lambda);

/* ---- ensemble runner ------------------------------------------------------ */
/* Serves independent systems of different sizes and seeds (the reference's
 * n_runs use case, src/simulation_manager.cpp:461-545, which runs them one
 * after another): requests are queued, admitted while the runner's device
 * memory (one arena, fixed at creation) and its concurrency limit allow,
 * kept resident on the device, advanced in chunks of steps, and retired --
 * final positions copied to the host, memory returned -- as they finish.
 * Admission is first-fit in arrival order.  Mixed-precision systems without
 * crosslinks, thermodynamic sums or a caller noise buffer are stepped in
 * batches: one set of launches per step for up to batch_size systems,
 * members joining and leaving at chunk boundaries; the others run their
 * chunks on their own streams.  A system's result does not depend on what
 * else runs beside it. */
typedef struct sc_ensemble_s *sc_ensemble;

typedef struct sc_ensemble_params {
  size_t hbm_budget;        /* arena bytes; 0: 80% of the free device memory */
  int32_t max_active;       /* systems resident at once; 0: 16             */
  int32_t steps_per_launch; /* steps per chunk (batches: at least 2); 0: 64 */
  int32_t batch_size;       /* systems per batch, at most 32; 0: 8; 1: no
                               batches (every system on its own stream)     */
} sc_ensemble_params;

/* request status */
#define SC_REQ_QUEUED 0
#define SC_REQ_RUNNING 1
#define SC_REQ_DONE 2
/* negative: failed with that SC_ERR_* code */

typedef struct sc_request_info {
  int32_t status;
  int32_t n_filaments;
  int64_t n_sites, n_bonds, n_steps, steps_done;
  double t_submit, t_admit, t_done; /* seconds on the ensemble clock       */
  size_t device_bytes;              /* footprint while resident (at the end
                                       for a finished request)             */
} sc_request_info;

typedef struct sc_ensemble_stats {
  int64_t submitted, completed, failed, running, queued;
  double latency_p50, latency_p99, latency_mean, latency_max; /* s, done  */
  double segment_steps;  /* bonds x steps of the completed requests        */
  double span;           /* first submission to last completion, s         */
  int32_t peak_active;   /* most systems resident at once                  */
  size_t peak_bytes;     /* most arena bytes in use at once                */
  size_t arena_bytes;
  int64_t batched_steps; /* system-steps taken in batches                  */
  int64_t solo_steps;    /* system-steps taken on the systems' own streams */
  /* host seconds spent in sc_ensemble_progress admitting systems, setting
   * up and enqueuing chunks, retiring finished systems */
  double host_admit_s, host_chunk_s, host_retire_s;
  double host_capture_s;  /* of host_chunk_s: capturing batch graphs      */
  int64_t graph_captures; /* batch graphs captured (membership changes)   */
} sc_ensemble_stats;

int sc_ensemble_create(const sc_ensemble_params *params, sc_ensemble *out);
/* Queues one system (inputs copied).  t_submit < 0: now. */
int sc_ensemble_submit(sc_ensemble e, const sc_params *params,
                       int32_t n_filaments, const int32_t *site_off,
                       const double *r, const sc_filament_params *fil,
                       int64_t n_steps, double t_submit, int64_t *request_id);
/* One scheduling round; never waits for the device. */
int sc_ensemble_progress(sc_ensemble e, int32_t *n_running, int32_t *n_queued);
/* Rounds until nothing is queued or running, or timeout_s (<= 0: none). */
int sc_ensemble_drain(sc_ensemble e, double timeout_s);
int sc_ensemble_request(sc_ensemble e, int64_t request_id, sc_request_info *info);
/* Final site positions (3*n_sites, SoA) of a finished request. */
int sc_ensemble_result(sc_ensemble e, int64_t request_id, double *r);
int sc_ensemble_stats_get(sc_ensemble e, sc_ensemble_stats *out);
/* Seconds since sc_ensemble_create (the clock of t_submit/t_admit/t_done). */
double sc_ensemble_clock(sc_ensemble e);
const char *sc_ensemble_last_error(sc_ensemble e);
int sc_ensemble_destroy(sc_ensemble e);

/* ---- host layer: parameter files and system generation ------------------- */

/* Reads a reference-style YAML parameter file (a subset of YAML: maps,
 * sequences of maps, flow lists, scalars, comments) merged over the
 * reference defaults (config/default_config.yaml) and fills sc_params for
 * the first filament species (src/params_parser.cpp:4-86,
 * src/simulation.cpp:248-251, src/interaction_manager.cpp:46-54).
 * Species values come back through `spec` (9 doubles, may be NULL): num,
 * length, persistence_length, diameter, min_bond_length, friction_ratio,
 * n_steps, min_length, max_length.  Only 3D, fully periodic, boundary 0,
 * WCA systems are accepted (SC_ERR_ARG otherwise). */
int sc_params_from_yaml(const char *path, sc_params *out, double *spec);

/* Fills sc_params with the reference defaults (config/default_config.yaml)
 * for a fully periodic cubic box of side 2*system_radius; the unit cell and
 * its inverse follow Space::InitUnitCell + invert_sym_3d_matrix
 * (src/space.cpp:85-138, src/linear_algebra.cpp:52-64). */
int sc_params_init(sc_params *out, double system_radius);

/* Filament::InitFilamentLength, static-length branch
 * (src/filament.cpp:104-147): bonds and bond length of a filament. */
int sc_filament_layout(double length, double min_bond_length,
                       double min_length, double max_length,
                       int32_t *n_bonds, double *bond_length);

/* Generates filament site positions in the reference's "random" insertion
 * style (Filament::InsertFilament / GenerateProbableOrientation,
 * src/filament.cpp:225-249, 350-383) with overlap rejection
 * (InteractionManager::CheckOverlap, src/interaction_manager.cpp:561-571),
 * using the build's own RNG.  Per filament: site count, persistence length,
 * diameter, total length.  Two-call pattern: with r == NULL only
 * *n_sites_total is written. */
int sc_generate_filaments(const sc_params *params, int32_t n_filaments,
                          const int32_t *n_sites, const double *length,
                          const double *persistence_length,
                          const double *diameter, double friction_ratio,
                          uint64_t seed, int32_t *site_off, double *r,
                          sc_filament_params *fil, int64_t *n_sites_total);

/* ---- trajectory files ------------------------------------------------------ */
/* The reference's binary .posit / .spec files for a filament species, from
 * caller arrays (host only; pair with sc_get_sites).  File names follow
 * SpeciesBase::InitPositFile: <run>_filament_<species>.posit / .spec.
 * sc_output_header creates the file with its header (int n_steps, int
 * n_posit or n_spec, double delta; src/species_base.cpp:13-27, 55-70);
 * sc_spec_frame / sc_posit_frame append one frame in the layout of
 * Species::WriteSpecs / WritePosits (include/simcore/species.hpp:221-236),
 * Filament::WriteSpec (src/mesh.cpp:482-492, src/filament.cpp:1303-1318)
 * and Filament::WritePosit (src/filament.cpp:1338-1352). */
int sc_output_header(const char *path, int32_t n_steps, int32_t n_every,
                     double delta);
int sc_spec_frame(const char *path, int32_t n_filaments, const int32_t *site_off,
                  const double *r, const sc_filament_params *fil);
int sc_posit_frame(const char *path, const sc_params *params,
                   int32_t n_filaments, const int32_t *site_off,
                   const double *r, const sc_filament_params *fil);

/* The reference's .thermo file (OutputManagerBase::InitThermo /
 * WriteThermo, include/simcore/output_manager.hpp:156-183): header int
 * n_steps, int n_thermo, double delta, int n_dim; per frame double[9] unit
 * cell, double[9] pressure tensor, double pressure, double volume (pair
 * with sc_calculate_pressure). */
int sc_thermo_header(const char *path, int32_t n_steps, int32_t n_thermo, double delta,
                     int32_t n_dim);
int sc_thermo_frame(const char *path, const double *unit_cell, const double *pressure_tensor,
                    double pressure, double volume);

#ifdef __cplusplus
}
#endif
#endif
