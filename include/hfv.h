/*
 * hfv — MI355X (gfx950) per-face finite-volume update for moving-mesh Voronoi hydrodynamics.
 *
 * C-ABI drop-in boundary for the hot path of the reference crate mvmm_hydro (yuyttenhove/hydro_rs):
 * the FiniteVolumeSolver trait (src/finite_volume_solver.rs:9-17), the RiemannFluxSolver trait
 * (src/riemann_solver.rs:78-87) and the runner functions that surround them
 * (src/runner/hydrodynamics.rs:13-195, src/runner/hydrodynamics/optimal_order.rs:15-64).
 *
 * Conventions
 *   - POD types only; no pointer is retained after a call returns (inputs are copied).
 *   - Device state lives in library-owned HBM buffers behind an opaque context; one context owns one
 *     HIP stream.  A context is not re-entrant; distinct contexts may be driven from distinct threads.
 *   - Return value: HFV_OK (0) or a negative error code.  Negative codes HFV_ERR_NONFINITE ..
 *     HFV_ERR_BRENT are the reference's hard assert!/panic! classes, raised by the device kernels;
 *     the reference's soft events (eprintln! + reset) are counted in hfv_stats.
 *   - Host arrays are structure-of-arrays, component-major: field[c * n + i].
 */
#ifndef HFV_H
#define HFV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- error codes */
#define HFV_OK 0
#define HFV_ERR_NONFINITE (-1)    /* src/part.rs:180-182, 202-204, 249-260 */
#define HFV_ERR_NEG_MASS (-2)     /* src/part.rs:220 */
#define HFV_ERR_NEG_ENERGY (-3)   /* src/part.rs:221-231 */
#define HFV_ERR_NEG_PRESSURE (-4) /* src/physical_quantities.rs:220-221, src/part.rs:246-247 */
#define HFV_ERR_NEWTON (-5)       /* src/riemann_solver/exact.rs:272-274 */
#define HFV_ERR_BRENT (-6)        /* src/riemann_solver/exact.rs:132-134 */
#define HFV_ERR_UNSUPPORTED (-7)
#define HFV_ERR_DRIFT (-8)        /* src/part.rs:118, src/utils.rs box_reflect containment */
#define HFV_ERR_INVALID (-20)     /* bad argument / shape */
#define HFV_ERR_HIP (-21)         /* HIP runtime error (hfv_last_error() has the text) */
#define HFV_ERR_MESH (-22)        /* tessellation failed */
#define HFV_ERR_NO_DEVICE (-23)

/* ---------------------------------------------------------------- enums */
/* Riemann solvers: src/riemann_solver.rs:28-44, src/bin/main.rs:625-642 */
#define HFV_EXACT 0
#define HFV_HLLC 1
#define HFV_PVRS 2
#define HFV_AIRS 3
#define HFV_TSRS 4
#define HFV_TRRS 5
#define HFV_LINEAR_ADVECTION 6
/* Boundary: src/space.rs:26-32 */
#define HFV_PERIODIC 0
#define HFV_REFLECTIVE 1
#define HFV_OPEN 2
#define HFV_VACUUM 3
/* ParticleMotion: src/engine.rs:7-13 */
#define HFV_FIXED 0
#define HFV_STEER 1
#define HFV_STEER_PAKMOR 2
#define HFV_FLUID 3
/* Runner: src/bin/main.rs:667-673 */
#define HFV_OPTIMAL_ORDER 0
#define HFV_GODUNOV 1
/* Equation of state: src/gas_law.rs:57-60 */
#define HFV_IDEAL 0
#define HFV_ISOTHERMAL 1
/* Arithmetic of the face reconstruction and Riemann solve. Accumulation, conserved->primitive and
 * the exact solver's pressure iteration stay FP64 in both modes.  FP32_FLUX contexts store gradients,
 * extrapolations and face records as floats and sum the engine steps' predicts where they are read
 * (rate x elapsed time): hfv_get_state reports the sums due at the current time, hfv_set_state and
 * the trait-level phases sum the pending part first, so results differ from per-step sums by FP32
 * rounding only. */
#define HFV_FP64 0
#define HFV_FP32_FLUX 1
/* Gravity solver (src/bin/main.rs:643-666): none, SelfGravity (src/gravity.rs:12-55), the external
 * constant potential and the external Keplerian potential (gravity.rs:57-149). */
#define HFV_GRAV_NONE 0
#define HFV_GRAV_SELF 1
#define HFV_GRAV_CONSTANT 2
#define HFV_GRAV_KEPLERIAN 3

/* ---------------------------------------------------------------- types */
/* EulerEqnsFvs::new + GasLaw + engine configuration (src/finite_volume_solver/euler_eqns_fvs.rs:7-17,
 * src/gas_law.rs:62-73, src/engine.rs:104-149, src/bin/main.rs:253-606). */
typedef struct {
  int dim;                 /* 1, 2 or 3 (Dimensionality) */
  int boundary;            /* HFV_PERIODIC .. HFV_VACUUM */
  double box[3];           /* box size (origin at 0) */
  double gamma;            /* adiabatic index */
  int eos;                 /* HFV_IDEAL / HFV_ISOTHERMAL */
  double isothermal_u;     /* isothermal internal energy */
  double cfl;              /* cfl_criterion */
  int solver;              /* HFV_EXACT .. HFV_LINEAR_ADVECTION */
  double airs_threshold;   /* AIRS p_max/p_min threshold */
  double linadv_velocity[3];
  int runner;              /* HFV_OPTIMAL_ORDER / HFV_GODUNOV */
  int motion;              /* HFV_FIXED .. HFV_FLUID */
  int sync_all;            /* time_integration:sync_timesteps */
  double t_end;
  double dt_snap;          /* dump sync points (drift only, no file output); <= 0: only t = 0 */
  int precision;           /* HFV_FP64 / HFV_FP32_FLUX */
  int gravity;             /* HFV_GRAV_NONE .. HFV_GRAV_KEPLERIAN (gravity: kind / potential: kind) */
  double grav_softening;   /* gravity:softening_length (self gravity, Keplerian potential) */
  double grav_position[3]; /* gravity:potential:position (Keplerian) */
  double grav_acceleration[3]; /* gravity:potential:acceleration (constant) */
} hfv_params;

/* Face-indexed mesh (host arrays) as produced by the tessellator; replaces the VoronoiFace slice and
 * cell->face connections of src/space.rs:76-77 (meshless_voronoi output). */
typedef struct {
  int64_t n_cells;
  int64_t n_faces;
  int64_t n_conn;             /* entries of conn[] */
  const double* area;         /* [F] */
  const double* normal;       /* [3F] xyz interleaved, unit, left -> right */
  const double* centroid;     /* [3F] in the left cell's frame */
  const double* shift;        /* [3F] periodic image offset of the right cell (0 when none) */
  const int64_t* left;        /* [F] */
  const int64_t* right;       /* [F], -1 for a wall face */
  const double* volume;       /* [N] */
  const double* cell_centroid;/* [3N] xyz interleaved */
  const double* search_radius;/* [N] */
  const int64_t* offset;      /* [N] face_connections_offset */
  const int64_t* count;       /* [N] face_count */
  const int64_t* conn;        /* [n_conn] cell -> face CSR (each cell's faces in accumulation order) */
} hfv_mesh_arrays;

/* Per-particle state of src/part.rs:10-39 in SoA form; host pointers, any may be NULL on set/get to
 * skip that field.  Shapes: prim/extrap/cons/fluxes [5N], grad [15N] ((comp*3+axis)*N + i),
 * gc/loc/centroid/v/v_rel [3N], scalars [N]. */
typedef struct {
  double* prim;
  double* grad;
  double* gc;       /* gradients_centroid */
  double* extrap;   /* extrapolations */
  double* cons;     /* conserved */
  double* fluxes;   /* pending flux accumulator */
  double* volume;
  double* loc;
  double* centroid;
  double* v;
  double* v_rel;
  double* maor;     /* max_a_over_r */
  double* msv;      /* max_signal_velocity */
  double* dt;
  int8_t* timebin;
} hfv_state;

/* TimestepInfo (src/engine.rs:33-81). */
typedef struct {
  uint64_t ti_old;
  uint64_t ti_current;
  uint64_t ti_next;
  double time_base;
  double time_base_inv;
} hfv_timestep_info;

/* Engine bookkeeping (src/engine.rs:84-101). */
typedef struct {
  hfv_timestep_info ts;
  double t_current;
  double t_end;
  int64_t step_count;
  int64_t cell_updates;    /* sum of active cells over all steps */
  int64_t face_fluxes;     /* sum of gated-in faces over all steps */
  int64_t last_n_active;
  int64_t last_n_face_active;
  int done;
  int error_code;          /* the simulation's first hard error (HFV_ERR_*), 0 if none; a batch slot
                              that fails stops (done = 1) while the other slots go on */
} hfv_engine_info;

/* Soft-event counters (the reference eprintln!s these and continues). */
typedef struct {
  int64_t check_physical_resets; /* src/physical_quantities.rs:245-253 */
  int64_t apply_flux_resets;     /* src/part.rs:184-191 */
  int64_t gradient_fallbacks;    /* src/gradients.rs:48-53 */
  int64_t brent_calls;           /* src/riemann_solver/exact.rs:279-284 */
  int64_t error_code;            /* first hard error raised on the device (0 = none) */
} hfv_stats;

typedef struct hfv_ctx hfv_ctx;

/* ---------------------------------------------------------------- lifecycle */
const char* hfv_last_error(void);
int hfv_device_count(int* n);
/* EulerEqnsFvs::new (euler_eqns_fvs.rs:13-17) + Engine::new (engine.rs:104-149). */
int hfv_create(const hfv_params* params, hfv_ctx** out);
int hfv_destroy(hfv_ctx* ctx);
/* Space::volume_calculation output (src/space.rs:236-263): copies the mesh to HBM. update_mask
 * (nullable, [N]) selects the cells whose volume/centroid/search radius are replaced
 * (Particle::update_geometry, src/part.rs:151-158); others only get new face offsets/counts.  The first
 * call gives every simulation its initial mesh; a simulation with moving generators then rebuilds its
 * mesh on the device at every step (hfv_step, hfv_run), so a mesh set between its steps is replaced by
 * the next step's rebuild. */
int hfv_set_mesh(hfv_ctx* ctx, const hfv_mesh_arrays* mesh, const uint8_t* update_mask);
/* Particle::from_ic + first_init (src/part.rs:200-217, src/space.rs:571-594): mesh must be set. */
int hfv_init_from_ic(hfv_ctx* ctx, int64_t n, const double* loc, const double* mass, const double* vel,
                     const double* u_int);
int hfv_set_state(hfv_ctx* ctx, const hfv_state* st);
int hfv_get_state(hfv_ctx* ctx, hfv_state* st);
int hfv_get_stats(hfv_ctx* ctx, hfv_stats* out);
int hfv_synchronize(hfv_ctx* ctx);

/* ---------------------------------------------------------------- FiniteVolumeSolver trait */
/* FiniteVolumeSolver::predict (src/finite_volume_solver.rs:10; euler_eqns_fvs.rs:20-24). */
int hfv_predict(hfv_ctx* ctx, double dt);
/* FiniteVolumeSolver::compute_fluxes (src/finite_volume_solver.rs:12; euler_eqns_fvs.rs:26-48).
 * The active mask is bin_is_ending(timebin) at info->ti_current (optimal_order.rs:17).
 * Flux records stay in HBM; hfv_get_fluxes copies them out as [7F]: dt*F (5), v_max, a_over_r.
 * FluxInfo.mflux (src/flux.rs:10-15, 96) is not computed: its only consumer is gravity_mflux in
 * grav_kick (part.rs:160-176, 303-312), where it is multiplied by a_grav, which the reference never
 * sets on this path (gravity is not wired into OptimalOrderRunner::step, optimal_order.rs:33). */
int hfv_compute_fluxes(hfv_ctx* ctx, const hfv_timestep_info* info);
int hfv_get_fluxes(hfv_ctx* ctx, double* out7f);
/* runner apply_fluxes + kick2 (src/runner/hydrodynamics.rs:13-57). */
int hfv_apply_fluxes(hfv_ctx* ctx);
/* FiniteVolumeSolver::convert_conserved_to_primitive (src/finite_volume_solver.rs:14). */
int hfv_conserved_to_primitive(hfv_ctx* ctx);
/* runner gradient_estimate/limit/apply (src/runner/hydrodynamics.rs:67-143). */
int hfv_gradients(hfv_ctx* ctx);
/* FiniteVolumeSolver::compute_timesteps + timesteps_apply + timestep_limiter/sync +
 * reset_extrapolations + kick1 (src/finite_volume_solver.rs:16; src/runner/hydrodynamics.rs:41-49,
 * 145-195); writes ti_current + min(dti) to *ti_next. */
int hfv_timesteps(hfv_ctx* ctx, uint64_t* ti_next);
/* FiniteVolumeSolver::compute_timesteps (src/finite_volume_solver.rs:16; euler_eqns_fvs.rs:50-61) on
 * its own: for the active cells Particle::timestep (part.rs:42-84) sets v and v_rel and resets
 * max_signal_velocity and max_a_over_r; dt_out [N] receives each particle's requested timestep, +inf
 * for inactive particles (the reference's Vec<f64>).  Nothing is applied: pass these (or the minimum
 * with other timesteps, optimal_order.rs:43-47) to hfv_apply_timesteps. */
int hfv_compute_timesteps(hfv_ctx* ctx, double* dt_out);
/* timesteps_apply + timestep_limiter/timestep_sync + reset_extrapolations + kick1
 * (src/runner/hydrodynamics.rs:145-195, optimal_order.rs:48-59) of the requested timesteps dt [N];
 * writes ti_current + min(dti) to *ti_next.  hfv_timesteps is hfv_compute_timesteps followed by this. */
int hfv_apply_timesteps(hfv_ctx* ctx, const double* dt, uint64_t* ti_next);

/* ---------------------------------------------------------------- GravitySolver trait */
/* GravitySolver::compute_timesteps (src/gravity.rs:6-10) for every particle, as the runner step sees it
 * (a_grav = 0): the Keplerian potential's eta * radius_2D / v_circ (gravity.rs:121-125, 140-148), +inf for
 * self gravity and the constant potential (gravity.rs:45-54, 98-103); dt_out [N].  The fused step
 * (hfv_step, hfv_run, batches) takes the minimum with the hydro timesteps itself (optimal_order.rs:47-50,
 * godunov.rs:41-44); with hfv_compute_timesteps / hfv_apply_timesteps the caller takes it.
 * HFV_ERR_INVALID when the context has no gravity solver (HFV_GRAV_NONE). */
int hfv_gravity_timesteps(hfv_ctx* ctx, double* dt_out);
/* GravitySolver::compute_accelerations (src/gravity.rs:6-10) at the current positions and masses: self
 * gravity's direct sum over the particles after each one (gravity.rs:24-44), the constant vector
 * (gravity.rs:82-89) or the softened Keplerian point mass (gravity.rs:127-138); a_out [3N] component-major.
 * The reference stores them in Particle::a_grav; this entry point only returns them.  No runner step calls
 * compute_accelerations (optimal_order.rs:33, godunov.rs:32), so on the step path a_grav stays zero and the
 * gravity kicks (hydrodynamics.rs:41-57) add nothing.  HFV_ERR_INVALID without a gravity solver. */
int hfv_gravity_accelerations(hfv_ctx* ctx, double* a_out);

/* ---------------------------------------------------------------- fused step and engine */
/* OptimalOrderRunner::step / GodunovHydroRunner::step (optimal_order.rs:15-64, godunov.rs:14-58).  For
 * moving meshes the step's Voronoi rebuild (Space::volume_calculation, space.rs:236-263; in 2D and 3D the
 * partial build of the step's active cells) runs on the device from the current generator positions. */
int hfv_step(hfv_ctx* ctx, const hfv_timestep_info* info, uint64_t* ti_next);
/* Engine::run (src/engine.rs:152-192): advance until max_steps Step sync points ran or t_end.  The
 * device loop drifts generators and rebuilds moving meshes.  A device rebuild that fails (more faces or
 * CSR entries than the context holds; a 2D cell of more than 48 edges; a 3D cell whose clipping takes more
 * than 64 planes (the box's 6 it starts from and every bisector that cuts it) or 98 vertices, or a face of
 * more than 24 vertices) stops the simulation with HFV_ERR_MESH. */
int hfv_run(hfv_ctx* ctx, int64_t max_steps, int64_t* steps_done);
int hfv_get_engine_info(hfv_ctx* ctx, hfv_engine_info* out);
/* Space::status sums (src/space.rs:1248-1280): sum of conserved + pending fluxes per component,
 * reduced on the device in a fixed order. */
int hfv_conservation(hfv_ctx* ctx, double* out5);
/* The current mesh in hfv_set_mesh's layout: Space's faces and cell -> face connections
 * (src/space.rs:76-77; Space::faces, space.rs:1232) and the cells' volume, centroid and search radius;
 * for moving meshes the device's latest rebuild.  With out->area == NULL only n_cells, n_faces and n_conn are written; then
 * call again with every array allocated at those sizes (they are written through the const pointers). */
int hfv_get_mesh(hfv_ctx* ctx, hfv_mesh_arrays* out);

/* ---------------------------------------------------------------- batched ensembles and serving */
/* Many independent simulations in one context.  No counterpart in the reference, which runs one
 * simulation per process (src/bin/main.rs:615-736); every simulation keeps the reference's semantics
 * (its own Engine, TimestepInfo and timebins), and one launch sequence, captured in a HIP graph,
 * advances every unfinished simulation by one of its own engine steps.
 *
 * A batch is a set of slots; slot s owns cell_cap[s] cells, face_cap[s] faces and conn_cap[s] CSR
 * entries of HBM.  A slot holds at most one simulation at a time; hfv_batch_load replaces it, so a
 * serving loop refills slots as simulations finish. */
int hfv_batch_create(int nslots, const int64_t* cell_cap, const int64_t* face_cap, const int64_t* conn_cap,
                     hfv_ctx** out);
/* Load a simulation into a slot: Space::initialize + first_init (src/space.rs:88-211) from the mesh and
 * the IC arrays (loc, vel [3n] xyz interleaved; mass, u_int [n]); resets the slot's engine (Engine::new,
 * src/engine.rs:104-149).  Fails with HFV_ERR_INVALID if the mesh exceeds the slot's capacity.  Slots of
 * moving simulations need face and CSR capacity above the initial mesh's (rebuilt meshes' counts vary). */
int hfv_batch_load(hfv_ctx* ctx, int slot, const hfv_params* params, const hfv_mesh_arrays* mesh,
                   const double* loc, const double* mass, const double* vel, const double* u_int);
/* Mark a slot empty (its cells stop participating). */
int hfv_batch_unload(hfv_ctx* ctx, int slot);
/* Enqueue up to max_launches batch steps; stops early (checked every few launches) when every loaded
 * simulation has finished.  *launches receives the number of batch steps enqueued. */
int hfv_batch_run(hfv_ctx* ctx, int64_t max_launches, int64_t* launches);
int hfv_batch_get_engine_info(hfv_ctx* ctx, int slot, hfv_engine_info* out);
int hfv_batch_get_state(hfv_ctx* ctx, int slot, hfv_state* st);
int hfv_batch_num_slots(hfv_ctx* ctx, int* n);

/* ---------------------------------------------------------------- geometry mode */
/* Fixed meshes (HFV_FIXED motion) whose particle velocities are all +0 use precomputed geometry: the
 * gradient weights and offsets per CSR entry, the inverse gradient matrix per cell, and the centroid
 * offset and a_over_r per face, evaluated once with the per-step expressions (no counterpart in the
 * reference, which recomputes them every step).  Results are bit-identical to use_static = 0.  Default
 * 1; hfv_get_geometry_mode reports whether a slot's simulation currently qualifies. */
int hfv_set_geometry_mode(hfv_ctx* ctx, int use_static);
int hfv_get_geometry_mode(hfv_ctx* ctx, int slot, int* is_static);

/* ---------------------------------------------------------------- autotuner */
/* Per size class of the live static-geometry simulations (one dimension, cell count in
 * [cells_lo, cells_hi); buckets 0 / 1024 / 8192 / 65536), hfv_autotune chooses the flux kernel's
 * tiling (its chunks' cells staged through LDS, or each face gathering its cells), the gradient
 * geometry (precomputed per CSR entry, or evaluated per step) and, in 3D, the gradient sweeps'
 * neighbour access (each cell chunk's neighbour lines staged through LDS, or gathered) by timing real launches
 * of each candidate in turn and keeping the fastest per cell update: per round and candidate, one
 * untimed launch, then groups of `launches` timed launches until they hold `launches` sweeps of the
 * live cells or 4 groups have run.  A class keeps its default (staged tiling in 2D and 3D; precomputed
 * geometry and staged neighbours in 3D) unless the fastest candidate beats it by more than 1%, the size of the
 * differences that the launches' changing mix of active cells produces on its own.  Every
 * candidate gives bit-identical results, so the tuning launches are ordinary steps of the run.  No
 * counterpart in the reference (it has no device kernels to tune).  ns_per_update[c]: measured cost of
 * candidate c = static_grad + 2 staged + 4 gstaged; -1 where not tried.  out may hold max_out classes;
 * n_out receives the number of classes. */
typedef struct {
  int dim;
  int n_sims;
  int64_t cells_lo, cells_hi;
  int staged;
  int static_grad;
  int gstaged;
  int pad_;
  double ns_per_update[8];
} hfv_tune_class;
int hfv_autotune(hfv_ctx* ctx, int launches, int rounds, hfv_tune_class* out, int max_out, int* n_out);
/* The same choices for one slot by hand: staged 1/0 (-1: by dimension, on for 2D and 3D, the
 * default), static_grad and gstaged 1/0 (-1: by dimension, on for 3D, the default).  They take effect for static-geometry
 * simulations only; hfv_get_tuning reports what a slot currently runs. */
int hfv_set_tuning(hfv_ctx* ctx, int slot, int staged, int static_grad, int gstaged);
int hfv_get_tuning(hfv_ctx* ctx, int slot, int* staged, int* static_grad, int* gstaged);
/* The flux kernel's launch shape for the whole context: face chunks per thread block, i.e. faces per
 * thread (1, 2 or 4; default 1): a block of 256 threads computes that many chunks of 256 faces in turn.
 * hfv_autotune, after the classes, times the three shapes the same way and keeps the fastest unless 1 is
 * within 1% of it; ns_per_update3 (nullable) receives its measured cost per cell update of each shape
 * (-1 where not tried).  Bit-identical results; no counterpart in the reference. */
int hfv_set_launch_shape(hfv_ctx* ctx, int flux_faces_per_thread);
int hfv_get_launch_shape(hfv_ctx* ctx, int* flux_faces_per_thread, double* ns_per_update3);

/* ---------------------------------------------------------------- measurement */
/* Device time of the step's kernels, from HIP events recorded on the context's stream around each
 * launch (no counterpart in the reference, which has no device).  Slots: 0 mask+predict, 1 mesh
 * rebuild (moving meshes), 2 face flux, 3 accumulate+c2p, 4 gradients+timestep, 5 limiter
 * wakeup, 6 limiter+kick1, 7 drift. */
#define HFV_NUM_KERNEL_SLOTS 8
typedef struct {
  int64_t launches[HFV_NUM_KERNEL_SLOTS];
  double ms[HFV_NUM_KERNEL_SLOTS];
} hfv_kernel_times;
/* enable != 0 starts recording (and clears the accumulators); 0 stops. */
int hfv_profile_enable(hfv_ctx* ctx, int enable);
int hfv_profile_read(hfv_ctx* ctx, hfv_kernel_times* out);

/* ---------------------------------------------------------------- unit entry point */
/* RiemannFluxSolver::solve_for_flux (src/riemann_solver.rs:78-87) over a batch of host arrays:
 * WL, WR [5n] primitive states (rho, vx, vy, vz, P), vface, nrm [3n], F [5n] out.  Component-major. */
int hfv_riemann_batch(const hfv_params* params, int64_t n, const double* WL, const double* WR,
                      const double* vface, const double* nrm, double* F);
/* hfv_riemann_batch without any device allocation: the same arguments and results (the same kernel),
 * staged through a fixed device array that the library reserves on each GPU it runs on (about 11 MB, held
 * until the process ends).  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (inside HIP graph capture, tight serving loops, allocation-free harnesses).  n is at most
 * HFV_NOALLOC_MAX_FACES per call (HFV_ERR_INVALID beyond it).  Every call on a GPU uses that GPU's one
 * array, so calls that overlap on one GPU (from several threads) race; calls on distinct GPUs do not. */
#define HFV_NOALLOC_MAX_FACES 65536
int hfv_riemann_batch_noalloc(const hfv_params* params, int64_t n, const double* WL, const double* WR,
                              const double* vface, const double* nrm, double* F);

#ifdef __cplusplus
}
#endif
#endif
