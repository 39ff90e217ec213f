/*
 * genesis_gpu.h — C ABI of the MI355X force-and-integration path.
 *
 * Plain C linkage, POD structs, pointers and sizes only.  Every entry point
 * returns an int status: 0 = ok, < 0 = error (see GS_E_*); gs_last_error()
 * gives the message of the last failing call on the calling thread.
 *
 * Ownership: device buffers belong to the handle; host buffers passed in are
 * caller-owned and copied.  One handle = one HIP stream.  Calls on one handle
 * must be serialised by the caller (the reference scheduler runs systems
 * sequentially, src/scheduler/mod.rs:83-89).  Results are deterministic run
 * to run: no floating-point atomics are used anywhere.
 *
 * Units and semantics follow the reference crate (ValentinVaninetti/genesis):
 * positions in simulation length units, epsilon in kelvin converted with
 * thermal_constant, the Berendsen tau in ticks.  Species are element indices
 * 0..9 in the order H, He, C, N, O, Na, Si, P, S, Fe
 * (src/physics/forces.rs:176-189).
 */
#ifndef GENESIS_GPU_H
#define GENESIS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_NSPECIES 10

/* status codes */
#define GS_OK 0
#define GS_E_INVALID (-1)     /* bad argument / validation failure           */
#define GS_E_HIP (-2)         /* HIP runtime error                           */
#define GS_E_STATE (-3)       /* call order error (e.g. no state uploaded)   */
#define GS_E_OVERFLOW (-4)    /* a device capacity was exceeded              */
#define GS_E_NONFINITE (-5)   /* a kernel saw a non-finite value             */

/* grid_semantics */
#define GS_GRID_REFERENCE 0   /* pair multiplicity of the reference grid: an
                                 axis with exactly 2 cells reports a pair once
                                 per wrapping offset (src/physics/grid.rs:134-151) */
#define GS_GRID_MINIMAGE 1    /* every unordered pair once (= brute force,
                                 src/physics/grid.rs:313-326)                 */

/*
 * Physics and schedule parameters.  Replaces the fields of Config that the
 * force pass, the integrator and the thermostat read
 * (src/config/mod.rs:17-260; ForceSystem::new src/systems/forces.rs:100-142;
 * ThermostatSystem::new src/systems/thermostat.rs:29-35).
 */
typedef struct gs_params {
    double box[3];                 /* universe.size (periodic torus extent)  */
    double dt;                     /* universe.dt                            */
    double sigma[GS_NSPECIES];     /* ElementTable sigma (after overrides)   */
    double epsilon_k[GS_NSPECIES]; /* ElementTable epsilon in kelvin         */
    double charge[GS_NSPECIES];    /* ElementTable charge (seeding default)  */
    double mass[GS_NSPECIES];      /* per-species Mass component value       */
    double thermal_constant;       /* physics.thermal_constant (k)           */
    double lj_cutoff_factor;       /* LJ_CUTOFF_FACTOR = 2.5                 */
    double coulomb_constant;       /* physics.coulomb_constant (k_e)         */
    double coulomb_cutoff;         /* physics.coulomb_cutoff (x max sigma)   */
    double gravity_constant;       /* physics.gravity_constant               */
    double gravity_cutoff;         /* physics.gravity_cutoff (x max sigma)   */
    double thermostat_temperature; /* physics.thermostat_temperature (K)     */
    double thermostat_tau;         /* physics.thermostat_tau (ticks)         */
    double bond_k_bind;            /* stats.bond_k_bind                      */
    double bond_min_periods;       /* stats.bond_min_periods                 */
    int32_t enable_boundaries;
    int32_t enable_electrostatics;
    int32_t enable_gravity;
    int32_t enable_thermostat;
    int32_t enable_bond_interaction;
    int32_t enable_bond_observation;
    int32_t grid_semantics;        /* GS_GRID_REFERENCE or GS_GRID_MINIMAGE  */
    int32_t reserved0;
} gs_params;

/* Per-tick observables (StatsSystem energy part, src/systems/aggregate.rs:86-121). */
typedef struct gs_observables {
    double kinetic;        /* K = sum 1/2 m |v|^2 after the thermostat           */
    double potential;      /* V (LJ + Coulomb + gravity) of this tick's forces   */
    double bond_energy;    /* V_bond of this tick's forces                       */
    double temperature;    /* (2/3)(K/N)/k                                        */
    double lambda;         /* Berendsen scale applied this tick (1 if off)       */
    double energy_total;   /* K + V + V_bond                                      */
} gs_observables;

/* FoF summary (ClusterStats, src/analysis/mod.rs:204-216). */
typedef struct gs_cluster_stats {
    int64_t monomers;
    int64_t aggregates;
    int64_t largest;
    double mean_size;
    int64_t bound_pairs;
} gs_cluster_stats;

/* Kernel timing of gs_step, measured with HIP events recorded on the
 * handle's stream around each phase of every tick (no reference
 * counterpart: the reference only times whole ticks, src/universe/mod.rs:143-158). */
typedef struct gs_timing {
    int64_t steps;          /* ticks timed since gs_timing_enable            */
    double step_ms;         /* sum over ticks of the whole tick              */
    double rebuild_ms;      /* kick/drift/wrap + cell-list rebuild           */
    double pair_ms;         /* pair kernel (the dominant kernel)             */
    double kick_ms;         /* second kick + kinetic reduction + Berendsen   */
    int64_t pairs_in_range; /* ordered pairs within the cutoff, last tick    */
} gs_timing;

/* Pair-path tuning (no reference counterpart: the reference has one CPU
 * path).  pair_mode 0 rebuilds the cell list every tick and evaluates the
 * 27-cell LDS tiles directly; pair_mode 1 builds, from the same tiles, a
 * Verlet list with cutoff rc + skin and reuses it while every atom stays
 * within skin/2 of its position at the build.  The device tests the
 * displacements every tick and rebuilds the lists itself when needed; the
 * host only schedules a full re-sort of the resident state every
 * sort_every ticks (memory locality).  pair_mode 2 keeps the same lists
 * per list-grid cell: one workgroup per cell stages the positions of the
 * cell's stencil candidates in LDS and evaluates its atoms' lists from
 * there (16-bit entries, lanes per atom chosen cell by cell; reported as
 * lanes_per_atom 0).  Every mode gives the reference's pair set. */
typedef struct gs_tuning {
    int32_t pair_mode;       /* 0 = cell tiles, 1 = Verlet list, 2 = cell-staged list */
    int32_t lanes_per_atom;  /* mode 1: 1, 2, 4, 8 or 16 lanes share a list           */
    double skin;             /* list mode: list cutoff = rc + skin                   */
    int64_t sort_every;      /* list mode: ticks between full re-sorts (0 = keep)    */
    int64_t list_builds;     /* list builds so far (read only)                       */
    int64_t list_replays;    /* calls replayed from their snapshot (read only)       */
    double ms_per_tick;      /* gs_autotune: time of the chosen setting              */
} gs_tuning;

typedef struct gs_universe gs_universe;

/* ---- lifecycle ------------------------------------------------------------ */

/* ForceSystem::new + ThermostatSystem::new (src/systems/forces.rs:100-142,
 * src/systems/thermostat.rs:29-35).  Allocates resident device state for
 * n_atoms atoms on the current HIP device. */
int gs_create(const gs_params* params, int64_t n_atoms, gs_universe** out);
int gs_destroy(gs_universe* u);
const char* gs_last_error(void);
/* Which HIP device the library runs on and its name (for logs). */
int gs_device_info(int* device, char* name, int name_len);

/* ---- state transfer (ECS columns, src/components/) ---------------------
 * SoA arrays of length n_atoms in entity (slot) order.  species[i] is the
 * element index 0..9; charge is the per-entity Charge component.  Any of
 * the acceleration pointers may be NULL on upload (then a = 0, which is the
 * state spawn_atom leaves, src/universe/mod.rs:272). */
int gs_upload_state(gs_universe* u,
                    const double* x, const double* y, const double* z,
                    const double* vx, const double* vy, const double* vz,
                    const double* ax, const double* ay, const double* az,
                    const uint8_t* species, const double* charge);
int gs_download_state(gs_universe* u,
                      double* x, double* y, double* z,
                      double* vx, double* vy, double* vz,
                      double* ax, double* ay, double* az);

/* ---- systems on the hot path (one call = one System::run) ------------------ */

/* ForceSystem::run (src/systems/forces.rs:163-353): rebuilds the cell list,
 * evaluates every pair within the cutoff, writes the Acceleration column and
 * returns PotentialEnergy / BondEnergy.  Any output pointer may be NULL. */
int gs_forces(gs_universe* u, double* potential, double* bond_energy);
/* Per-atom extras of the last gs_forces call, in slot order: per-atom
 * potential energy (1/2 sum_j V_ij) and S_i = sum_j |f_ij|.  NULL skips. */
int gs_download_force_extras(gs_universe* u, double* pe_atom, double* s_abs);
/* Cell membership of the force grid for the current positions, in slot order:
 * cell = ix + nx*(iy + ny*iz) with (ix, iy, iz) from SpatialGrid::cell_of and
 * encode (src/physics/grid.rs:228-246) on a grid of edge >= the force cutoff
 * (ForceSystem::new, src/systems/forces.rs:121-124).  dims (u32[3]) may be NULL. */
int gs_download_cells(gs_universe* u, uint32_t* cell, uint32_t* dims);
/* VelocityHalfKick::run (src/systems/integrate.rs:38-44). */
int gs_half_kick(gs_universe* u);
/* PositionDrift::run (src/systems/integrate.rs:61-67). */
int gs_drift(gs_universe* u);
/* BoundarySystem::run (src/systems/boundary.rs:22-35). */
int gs_wrap(gs_universe* u);
/* ThermostatSystem::run (src/systems/thermostat.rs:47-71); lambda may be NULL. */
int gs_thermostat(gs_universe* u, double* lambda);
/* Kinetic energy sum (src/systems/aggregate.rs:86-93). */
int gs_kinetic(gs_universe* u, double* kinetic);
/* StatsSystem's reductions over Velocity and Mass (src/systems/aggregate.rs:
 * 87-93): kinetic = sum 0.5 m |v|^2 and speed_sum = sum |v| over all
 * entities, each in a fixed order (the reference sums sequentially; the two
 * agree to rounding).  mean_speed = speed_sum / n. */
int gs_speed_stats(gs_universe* u, double* kinetic, double* speed_sum);

/* Universe::tick x n_steps (src/universe/mod.rs:141-159) with the schedule of
 * build_schedule (src/universe/mod.rs:355-392): kick, drift, [wrap], forces,
 * kick, [thermostat].  Fused and resident: nothing crosses to the host but the
 * observables.  obs may be NULL, else it receives n_steps records.
 * The call runs in segments of 1024 ticks, each with its own replay
 * snapshot: a list or bucket that overflows replays only its segment.
 * With bond observation, an entity with more than 32 bound partners, tracked
 * episodes or bonds (the reference has no cap) makes the call return
 * GS_E_OVERFLOW; on the Verlet-list path the universe is then back at the
 * start of the segment that overflowed (gs_get_tick tells which), on the
 * cell tiles it holds that segment's end.  The overflow is sticky until
 * gs_bond_reset.
 * Device fault word: the last workgroup of every tick marks a non-finite
 * kinetic, potential or bond energy (what a NaN or infinite velocity, force
 * or position turns into within its tick); a call that saw one returns
 * GS_E_NONFINITE naming the quantities and the segment's ticks, with the
 * state left at that segment's end.  The reference has no such check. */
int gs_step(gs_universe* u, int64_t n_steps, gs_observables* obs);

int gs_get_tuning(gs_universe* u, gs_tuning* out);
/* Validates and applies; GS_E_INVALID if the setting is impossible for this
 * box (list mode needs >= 3 cells of edge rc + skin per axis) or universe
 * (the Verlet list, pair_mode 1, holds at most 2^27 atoms). */
int gs_set_tuning(gs_universe* u, const gs_tuning* t);
/* No counterpart in the reference (where gs_step finds the bound pairs of
 * bond observation).  0 (default): on the Verlet-list path with 2 or 4 lanes
 * per atom and k_bind max(sigma) <= rc, the pair kernel screens its list
 * entries against the bind cutoff and applies bound_pairs_with_grid's exact
 * test to the hits (src/analysis/pairs.rs:62-85); otherwise, and in a call
 * replayed because an atom had more candidates than that search holds, the
 * aux-grid search runs.  1: always the aux-grid search.  Results are
 * identical either way. */
int gs_set_bound_search(gs_universe* u, int32_t aux_grid_only);
/* Project autotuner: times trial_steps ticks of every candidate setting on
 * the resident state, restores the state after each trial and keeps the
 * fastest.  best (may be NULL) receives it.  A trial that fails (other than
 * with GS_E_HIP) is skipped; if every trial fails, the state, the tuning and
 * the timing mode are left as they were and the last trial's status is
 * returned.  The trials observe bonds too; the bond observation state
 * (tracker, adjacency, tick counter) is reset afterwards, as by
 * gs_bond_reset. */
int gs_autotune(gs_universe* u, int32_t trial_steps, gs_tuning* best);

/* Event timing of gs_step, reset on every call: on = 0 off, 1 the pair
 * kernel only (two events per timed tick), 2 every phase (four events per
 * timed tick; rebuild_ms and kick_ms stay 0 in mode 1).  Every `every`-th
 * tick of a gs_step call is timed (tick 0, every, 2 every, ...), which
 * bounds the cost of the event records; gs_timing.steps counts the timed
 * ticks. */
int gs_timing_enable(gs_universe* u, int on, int every);
int gs_timing_read(gs_universe* u, gs_timing* out);

/* ---- observation lenses ---------------------------------------------------- */

/* velocity_histogram (src/stats/mod.rs:255-297): speeds |v| of every atom in
 * nbins (clamped to 1..512) bins of width max(max_speed, 1e-9) / nbins;
 * counts: u64[nbins]; speeds >= max_speed count as overflow.  Any output may
 * be NULL. */
int gs_velocity_histogram(gs_universe* u, double max_speed, int32_t nbins, uint64_t* counts,
                          uint64_t* samples, uint64_t* overflow, double* bin_width);
/* radial_distribution (src/analysis/mod.rs:87-133), incl. the deterministic
 * subsample above 8192 atoms (:28-46).  counts: u64[nbins] pair counts,
 * g: f64[nbins]; r_max_out/dr_out receive the clipped r_max and bin width.
 * Any output may be NULL. */
int gs_radial_distribution(gs_universe* u, double r_max, int32_t nbins,
                           uint64_t* counts, double* g,
                           double* r_max_out, double* dr_out);
/* g(r) over every atom, without the subsample (no counterpart in the
 * reference, whose radial_distribution subsamples above 8192 atoms,
 * src/analysis/mod.rs:28-46): the same bins and normalisation with all n
 * atoms.  species_a < 0 gives the total g(r), otherwise g_ab(r) as
 * gs_radial_distribution_between. */
int gs_radial_distribution_full(gs_universe* u, int32_t species_a, int32_t species_b, double r_max,
                                int32_t nbins, uint64_t* counts, double* g, double* r_max_out, double* dr_out);
/* No counterpart in the reference: an allocation-free variant of
 * gs_radial_distribution for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same subsample, grid and pair counts (bit-identical to
 * gs_radial_distribution's counts); the counts are written to counts_dev, a
 * DEVICE array of nbins u64 (nbins in [1, 65536], otherwise GS_E_INVALID),
 * on the handle's stream, and the call does not wait for them: they are
 * ready once the stream is (any later call on the handle that returns data,
 * or a device synchronisation).  The subsample (at most 8192 points) and its
 * grid (at most 36864 cells) live in one fixed scratch area of the library
 * per GPU, under 1 MiB, held until the process ends; the call makes no
 * hipMalloc/hipFree and no host allocation.  Calls on one GPU share that
 * area: calls that overlap on one GPU (two threads, two handles' streams)
 * race.  r_max_out / dr_out (host, may be NULL) as in gs_radial_distribution;
 * g is the caller's normalisation of the counts. */
int gs_radial_distribution_noalloc(gs_universe* u, double r_max, int32_t nbins, uint64_t* counts_dev,
                                   double* r_max_out, double* dr_out);
/* radial_distribution_between (src/analysis/mod.rs:143-201). */
int gs_radial_distribution_between(gs_universe* u, int32_t species_a, int32_t species_b,
                                   double r_max, int32_t nbins,
                                   uint64_t* counts, double* g,
                                   double* r_max_out, double* dr_out);
/* clusters (src/analysis/mod.rs:221-284): friends-of-friends with
 * d < bond_factor * (sigma_a + sigma_b)/2.  labels (u32[n], may be NULL) get
 * the canonical component label = smallest slot index in the component. */
int gs_clusters(gs_universe* u, double bond_factor, uint32_t* labels,
                gs_cluster_stats* stats);

/* ---- observed bonds ---------------------------------------------------------
 * With enable_bond_observation, every gs_step tick ends with
 * BondObservationSystem::run (src/systems/bonds.rs:73-185): the pairs with
 * |min_image(a - b)| < bond_k_bind * sigma_ab (bound_pairs_with_grid,
 * src/analysis/pairs.rs:132-174) feed a PairTracker with a 2-tick debounce
 * (:40-111); an episode that reaches ceil(bond_min_periods * T_vib / dt)
 * ticks is a persistent bond and enters the Bonds adjacency, which the
 * force pass of the next tick uses when enable_bond_interaction is on (the
 * harmonic term of src/systems/forces.rs:305-317).  Entity ids are the
 * positions in the arrays given to gs_upload_state.  The tracker, the
 * adjacency and the tick counter survive gs_upload_state (a new state is a
 * new frame of the same world); gs_bond_reset clears them. */
typedef struct gs_bond_stats {    /* BondObservation (src/systems/bonds.rs:140-184) */
    int64_t tick;                 /* tick of the last observation                */
    int64_t bonded_pairs;         /* persistent pairs now                        */
    int64_t bonded_entities;      /* entities with at least one persistent bond  */
    double mean_coordination;     /* 2 bonded_pairs / bonded_entities            */
    int64_t bonds_formed;         /* persistent pairs that broke, cumulative     */
    double lifetime_sum_ticks;    /* their summed persistence, cumulative        */
    int64_t species_matrix[GS_NSPECIES * GS_NSPECIES]; /* [a*10+b] and [b*10+a] */
} gs_bond_stats;

/* BondObservationSystem::run on the resident positions as tick `tick`
 * (ctx.time.tick); later gs_step ticks continue from tick + 1. */
int gs_bond_observe(gs_universe* u, int64_t tick);
int gs_bond_stats_read(gs_universe* u, gs_bond_stats* out);
/* The Bonds adjacency as CSR over entity ids: ptr[n+1], idx[cap] with each
 * entity's neighbours ascending; *nnz receives the entry count
 * (GS_E_OVERFLOW when it exceeds cap). */
int gs_download_bonds(gs_universe* u, int32_t* ptr, int32_t* idx, int64_t cap, int64_t* nnz);
/* Replaces the Bonds adjacency (CSR over entity ids, made symmetric and
 * sorted; at most 32 neighbours per entity); the tracker is untouched. */
int gs_upload_bonds(gs_universe* u, const int32_t* ptr, const int32_t* idx);
/* bound_pairs_with_grid on the resident positions: pairs (a < b, entity
 * ids) ascending; *count may exceed cap (then only cap are written). */
int gs_bound_pairs(gs_universe* u, uint32_t* a, uint32_t* b, int64_t cap, int64_t* count);
/* Connected components of the Bonds graph (bond_components,
 * src/analysis/mod.rs:337-412): labels[n] by entity id = smallest entity
 * id of its component, 0xFFFFFFFF for entities without bonds. */
int gs_bond_components(gs_universe* u, uint32_t* labels, int64_t* n_components);
/* Observed binding energy of the Bonds graph on the resident positions
 * (BondStructureSystem::run, src/systems/bond_structure.rs:95-123, with
 * pair_potential_raw, src/physics/forces.rs:360-384): binding[e] (by entity
 * id) = sum over e's bonded partners m > e of the raw, unswitched LJ term
 * plus, when enabled, the Coulomb term with the species charges and the
 * harmonic bond term, at r = max(|min_image(x_e - x_m)|, 1e-9).  A
 * component's binding energy is the sum over its members. */
int gs_bond_binding(gs_universe* u, double* binding);
/* Clears the tracker, the Bonds adjacency, the cumulative counters and the
 * tick counter. */
int gs_bond_reset(gs_universe* u);

/* ---- snapshots (save_universe / load_universe, src/serialization/mod.rs:84-120)
 * The file format is the reference's bincode v4 Snapshot (config, clock,
 * RNG, stats and every entity's components); the host mirror writes and
 * reads that part (genesis_amd/snapshot.py).  These two calls add what the
 * device holds beyond the components, appended after the Snapshot (the
 * reference's loader ignores trailing bytes): the resident order, the
 * Verlet lists with their build positions, buckets and flags, the bond
 * tracker's episodes and counters, and the tick.  A universe created with
 * the same gs_params and atom count and loaded from the image steps bit for
 * bit like the one it was taken from on the per-tick cell tiles and the
 * Verlet lists (pair_mode 0 and 1); the cell-staged lists of pair_mode 2 are
 * not in the image, so such a universe re-sorts and rebuilds them on its
 * first tick and continues to within rounding.  Paths are plain file names.
 * bincode 1 (Cargo.toml: bincode = "1") reads a Snapshot with
 * `deserialize`, which allows trailing bytes. */
/* Appends the device image of u to the file (created if absent). */
int gs_snapshot_append(gs_universe* u, const char* path);
/* Restores the device image found at byte `offset` of the file; GS_E_INVALID
 * if it was taken from a universe with other parameters or atom count. */
int gs_snapshot_load(gs_universe* u, const char* path, int64_t offset);
/* Time::tick of the universe (restored by load_universe from a file without
 * a device image, src/serialization/mod.rs:116 and src/universe/mod.rs:129):
 * the next gs_step tick is tick + 1.  0 <= tick < 2^32 - 1. */
int gs_set_tick(gs_universe* u, int64_t tick);
/* Time::tick of the universe: the last tick gs_step ran (or the one set). */
int gs_get_tick(gs_universe* u, int64_t* tick);

/* ---- differentiable trajectory (no reference counterpart) -------------------
 * theta[30] = per-species sigma[10], epsilon_k[10], charge[10]; keep[n] (u8,
 * slot order) multiplies each atom's charge (0 = neutralised atom).  These
 * calls replace the handle's sigma/epsilon/charge by theta's, switch the
 * thermostat off and use the per-tick cell-tile path (deterministic, so a
 * recomputed segment reproduces the forward bit for bit).  LJ + Coulomb with
 * the Coulomb cutoff equal to the LJ cutoff; gravity and bonds off. */
typedef struct gs_gr_spec {
    int32_t nbins;    /* Gaussian-smoothed g(r): bins on [0, r_max), 1..8192 */
    int32_t pad;
    double r_max;
    double sigma_b;   /* Gaussian width (same length unit); the kernel is cut
                         where it is below 1.3e-14 of its peak: a pair feeds
                         the bins whose centre lies within 8 sigma_b of its
                         distance                                          */
} gs_gr_spec;

/* One adjoint pass at the resident positions: with u = lam / m,
 * out = (da/dx)^T lam - vcoef dV/dx  (a Hessian-vector product) and
 * dtheta = (da/dtheta)^T lam + vcoef dV/dtheta.  lam, out in slot order. */
int gs_hvp(gs_universe* u, const double* theta, const uint8_t* keep,
           const double* lx, const double* ly, const double* lz, double vcoef,
           double* ox, double* oy, double* oz, double* dtheta);
/* Forward op: n_steps velocity-Verlet ticks (NVE, a_0 = a(x_0)) from the
 * resident state, checkpoints every ckpt_every ticks, loss = mean V over
 * ticks [win_from, n_steps] + sum_k (g_k(x_T) - g0_k)^2 (gr and g0 may be
 * NULL).  g_final (nbins, may be NULL) gets g(x_T).  Leaves the state at x_T. */
int gs_adjoint_forward(gs_universe* u, const double* theta, const uint8_t* keep,
                       int64_t n_steps, int64_t ckpt_every, int64_t win_from,
                       const gs_gr_spec* gr, const double* g0,
                       double* loss, double* g_final);
/* Backward op: dloss/dtheta (30) of the last forward, by the discrete adjoint
 * with recomputation from the checkpoints. */
int gs_adjoint_backward(gs_universe* u, double* dloss_dtheta);

/* ---- batched replicas (no reference counterpart; replaces the sequential
 * Universe-per-temperature loop of examples/temperature_sweep.rs:70-113) -----
 * R independent universes of uneven sizes stepped together: one launch per
 * phase for all replicas.  params[r] and n_atoms[r] describe replica r; all
 * replicas share the element table, the enabled laws and cutoffs, and may
 * differ in box, dt and thermostat settings.  The pair path is chosen by
 * gs_batch_set_tuning: Verlet lists (pair_mode 1, the default: 2 lanes per
 * atom, skin 0.12 rc, re-sort every 100 ticks) for every replica with >= 3
 * list cells of edge rc + skin per axis, or the per-tick cell tiles
 * (pair_mode 0) for every replica with >= 3 cells per axis.  The other
 * replicas use a brute-force pair kernel (grid_semantics as in gs_params).
 * Bond observation and the bond interaction are not part of the batched
 * path: gs_batch_create rejects params with either enabled (GS_E_INVALID). */
typedef struct gs_batch gs_batch;
/* (at most 2^27 atoms over all replicas) */
int gs_batch_create(int32_t n_replicas, const gs_params* params, const int64_t* n_atoms, gs_batch** out);
int gs_batch_destroy(gs_batch* b);
/* State of replica r (slot order within the replica); every replica must be
 * uploaded once before gs_batch_step. */
int gs_batch_upload(gs_batch* b, int32_t replica,
                    const double* x, const double* y, const double* z,
                    const double* vx, const double* vy, const double* vz,
                    const double* ax, const double* ay, const double* az,
                    const uint8_t* species, const double* charge);
int gs_batch_download(gs_batch* b, int32_t replica,
                      double* x, double* y, double* z,
                      double* vx, double* vy, double* vz,
                      double* ax, double* ay, double* az);
/* n_steps ticks of every replica; graph_segment > 0 captures that many ticks
 * (rounded up to even) in a hipGraph once and replays it.  obs (may be NULL)
 * receives n_steps * n_replicas records, tick-major. */
int gs_batch_step(gs_batch* b, int64_t n_steps, int32_t graph_segment, gs_observables* obs);
/* Project autotuner of the batched path: the cell tiles with cell edge
 * s * rc, s in {1, 1.15, 1.3}, and the Verlet lists with skin in
 * {0.06, 0.12, 0.18} rc at 2 and 4 lanes per atom; keeps the fastest (read
 * it with gs_batch_get_tuning).  best_scale: the tiles' cell scale of the
 * kept setting (1 for the lists).  Outputs may be NULL. */
int gs_batch_autotune(gs_batch* b, int32_t trial_steps, double* best_scale, double* best_ms_per_tick);
/* Pair-path tuning of the batch (gs_tuning as for one universe; pair_mode 0
 * or 1; skin in length units; sort_every 0 keeps the current period).  The
 * setting applies to every replica whose box admits it. */
int gs_batch_get_tuning(gs_batch* b, gs_tuning* out);
int gs_batch_set_tuning(gs_batch* b, const gs_tuning* t);

#ifdef __cplusplus
}
#endif

#endif /* GENESIS_GPU_H */
