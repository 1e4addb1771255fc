/*
 * CPU FP64 oracle: a plain-C restatement of the reference crate `mvmm_hydro`
 * (yuyttenhove/hydro_rs) along the per-face finite-volume update.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the reference
 * that results are compared against or as the timed CPU baseline.  The product (hydro_rs_amd) never links it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * written in the reference's evaluation order (Rust evaluates left to right,
 * no FMA contraction); the library is compiled with -ffp-contract=off and
 * glibc libm, matching Rust's f64 semantics on linux-gnu.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - exact solver: 6 golden half-states (src/riemann_solver/exact.rs:363-397),
 *     symmetry, Galilean invariance, SWIFT near-vacuum regression, Toro table.
 *   - vacuum solver: golden half-state (src/riemann_solver/vacuum.rs:211-219).
 *   - HLLC / PVRS / AIRS / TSRS / TRRS: symmetry/invariance tests of the
 *     reference, consistency F(W,W) = Euler flux, HLLC within 5e-2 of exact.
 *   - AIRS: bitwise PVRS below the pressure-ratio threshold, exact at or above it.
 *   - pairwise limiter: hand-computed clamps; monotone, idempotent, scale/shift invariant.
 *   - WLS gradients: exact for linear fields; LimiterData alpha in closed form on a
 *     uniform 1D mesh (linear, extremum, one-sided, flat).
 *   - c2p: the reference's round trip (physical_quantities.rs:318-339, 4 ulp) and
 *     random round trips through the textbook p2c.
 *   - time_extrapolations: the conservative Euler flux derivative (1D), rotation and
 *     moving-frame invariance.
 *   - radius / Particle::timestep: closed forms; timestep_limit: brute force against
 *     the particle's time line; face velocity: the moving bisector plane.
 *   - timeline: brute force over all bins, round trip test (timeline.rs:149-155).
 *   - step: conservation of mass/momentum/energy to roundoff.
 *   - 1D tessellation: test_init_1d known answers (src/space.rs:1453-1491).
 * Parity unpinned: glam's exact DMat3::inverse operation order (restated from
 * glam 0.27's published source semantics; no binary to compare against).
 */
#ifndef HYDRO_ORACLE_H
#define HYDRO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Riemann solver kinds (src/riemann_solver.rs:28-44, src/bin/main.rs:625-642) */
enum { ORA_EXACT = 0, ORA_HLLC = 1, ORA_PVRS = 2, ORA_AIRS = 3, ORA_TSRS = 4, ORA_TRRS = 5, ORA_LINADV = 6 };
/* Boundary (src/space.rs:26-32) */
enum { ORA_PERIODIC = 0, ORA_REFLECTIVE = 1, ORA_OPEN = 2, ORA_VACUUM = 3 };
/* ParticleMotion (src/engine.rs:7-13) */
enum { ORA_FIXED = 0, ORA_STEER = 1, ORA_STEER_PAKMOR = 2, ORA_FLUID = 3 };
/* Runner (src/bin/main.rs:667-673) */
enum { ORA_OPTIMAL_ORDER = 0, ORA_GODUNOV = 1 };
/* Equation of state (src/gas_law.rs:57-60) */
enum { ORA_IDEAL = 0, ORA_ISOTHERMAL = 1 };

/* Error codes: the reference's hard assert!/panic! classes. */
enum {
  ORA_OK = 0,
  ORA_ERR_NONFINITE = -1,       /* part.rs:180-182, 202-204, 249-260 */
  ORA_ERR_NEG_MASS = -2,        /* part.rs:220 */
  ORA_ERR_NEG_ENERGY = -3,      /* part.rs:221-224, 226-231 */
  ORA_ERR_NEG_PRESSURE = -4,    /* physical_quantities.rs:220-221, part.rs:246-247 */
  ORA_ERR_NEWTON = -5,          /* exact.rs:272-274 */
  ORA_ERR_BRENT = -6,           /* exact.rs:132-134 */
  ORA_ERR_UNSUPPORTED = -7,
  ORA_ERR_DRIFT = -8,           /* part.rs:118, space.rs box_reflect assert */
  ORA_ERR_MESH = -22,           /* the tessellation of a 2D/3D moving mesh failed */
};

/* Soft-event counters (the reference eprintln!s these). */
typedef struct {
  int64_t check_physical_resets;  /* physical_quantities.rs:245-253 */
  int64_t apply_flux_resets;      /* part.rs:184-191 */
  int64_t gradient_fallbacks;     /* gradients.rs:48-53 */
  int64_t brent_calls;            /* exact.rs:279-284 */
} ora_counters;

typedef struct {
  int kind;
  double threshold;   /* AIRS */
  double velocity[3]; /* LinearAdvection */
} ora_riemann;

/* ---- Riemann solver unit API ---- */
/* RiemannFluxSolver::solve_for_flux (src/riemann_solver.rs:78-87). left/right: rho,vx,vy,vz,P */
int ora_solve_for_flux(const ora_riemann* rs, double gamma, const double* left, const double* right,
                       const double* iface_v, const double* n_unit, double* flux_out);
/* RiemannStarSolver::sample (src/riemann_solver.rs:224-255) (kind VACUUM=100 uses vacuum.rs:116-138) */
int ora_sample(int kind, double threshold, double gamma, const double* left, const double* right, double v_l,
               double v_r, double a_l, double a_r, const double* n_unit, double* half_out);
/* ExactRiemannSolver::solve_for_star_state (src/riemann_solver/exact.rs:242-290): out = rho_l, rho_r, u, p */
int ora_exact_star(double gamma, const double* left, const double* right, double v_l, double v_r, double a_l,
                   double a_r, double* star_out);
/* pairwise_limiter (src/gradients.rs:134-152) */
void ora_pairwise_limiter(const double* pl, const double* pr, const double* pdash, double dx_fac, double* out);
/* GasLaw::sound_speed (src/gas_law.rs:99-106) */
double ora_sound_speed(double gamma, double pressure, double density_inv);

/* ---- timeline (src/timeline.rs) ---- */
uint64_t ora_get_integer_timestep(int8_t bin);
int8_t ora_get_time_bin(uint64_t dti);
uint64_t ora_get_integer_time_end(uint64_t ti_current, int8_t bin);
int8_t ora_get_max_active_bin(uint64_t ti);
uint64_t ora_make_integer_timestep(double dt, int8_t old_bin, int8_t min_ngb_bin, uint64_t ti_current,
                                   double time_base_inv);
uint64_t ora_saturating_f64_to_u64(double x);

/* ---- simulation ---- */
typedef struct ora_sim ora_sim;

typedef struct {
  int dim;          /* 1, 2, 3 */
  int boundary;     /* ORA_* boundary */
  double box[3];
  double gamma;
  int eos;
  double isothermal_u;
  double cfl;
  ora_riemann riemann;
  int runner;
  int motion;
  int sync_all;
  double t_end;
  double dt_snap;   /* dump sync points are drifted to (no file I/O); <=0 => none besides t=0 */
  int gravity;      /* ORA_GRAV_* (src/gravity.rs) */
  double grav_softening;
  double grav_position[3];      /* Keplerian potential centre */
  double grav_acceleration[3];  /* constant potential */
} ora_params;
/* GravitySolver kinds (src/gravity.rs): none, SelfGravity, ExternalPotentialGravity(Constant|Keplerian) */
enum { ORA_GRAV_NONE = 0, ORA_GRAV_SELF = 1, ORA_GRAV_CONSTANT = 2, ORA_GRAV_KEPLERIAN = 3 };

/* Mesh as produced by the tessellator (replaces meshless_voronoi output, src/space.rs:236-263). */
typedef struct {
  int64_t n_faces;
  const double* area;      /* [F] */
  const double* normal;    /* [3F] xyz interleaved */
  const double* centroid;  /* [3F] */
  const int64_t* left;     /* [F] */
  const int64_t* right;    /* [F], -1 = boundary */
  const double* shift;     /* [3F], zero when the face has no periodic shift */
  int64_t n_conn;
  const int64_t* conn;     /* cell->face CSR entries */
  const int64_t* offset;   /* [N] */
  const int64_t* count;    /* [N] */
  const double* volume;    /* [N] */
  const double* cell_centroid; /* [3N] */
  const double* search_radius; /* [N] (already 1.5 * safety radius) */
} ora_mesh;

/* Space::initialize + first_init_parts (src/space.rs:88-211, 571-594) from IC arrays
 * (Particle::from_ic, src/part.rs:207-217) and a mesh. loc: [3N]. */
ora_sim* ora_sim_new(const ora_params* p, int64_t n, const double* loc, const double* mass, const double* vel,
                     const double* u_int, const ora_mesh* mesh);
void ora_sim_free(ora_sim* s);
/* 2D/3D moving meshes: Space::volume_calculation (space.rs:236-263) calls meshless_voronoi's
 * Voronoi::build_partial, which the oracle does not restate.  The step calls `mesher` with the generator
 * positions (loc [3N], xyz interleaved) and the active mask [N] after the predict; the mesher tessellates
 * them (the cells of the active generators) and hands the mesh to ora_sim_apply_mesh, which replaces the
 * faces and connections and updates the geometry of the step's active cells only.  A nonzero return stops
 * the simulation with ORA_ERR_MESH. */
typedef int (*ora_mesher)(void* user, ora_sim* s, int64_t n, const double* loc, const uint8_t* mask);
void ora_sim_set_mesher(ora_sim* s, ora_mesher mesher, void* user);
void ora_sim_apply_mesh(ora_sim* s, const ora_mesh* mesh);
int ora_sim_set_threads(int n);

/* Engine::run (src/engine.rs:152-192): process sync points until `max_steps` Step sync
 * points ran or t_end is reached. Returns the number of steps done, or an error code < 0. */
int64_t ora_sim_advance(ora_sim* s, int64_t max_steps);
/* Run one OptimalOrderRunner::step / GodunovHydroRunner::step at the current sync point without
 * the engine bookkeeping; returns ti_next via out. */
int ora_sim_runner_step(ora_sim* s, uint64_t* ti_next);

/* Engine state */
typedef struct {
  uint64_t ti_old, ti_current, ti_next;
  double time_base, time_base_inv, t_current, t_end;
  int64_t step_count;
  int error;
  int64_t last_n_active;
  int64_t last_n_face_active;
} ora_engine_info;
void ora_sim_engine_info(const ora_sim* s, ora_engine_info* out);
void ora_sim_set_engine_info(ora_sim* s, uint64_t ti_old, uint64_t ti_current, uint64_t ti_next);
void ora_sim_counters(const ora_sim* s, ora_counters* out);

/* Particle field exchange in the SoA layout of include/hfv.h (component-major). */
typedef struct {
  double* prim;     /* [5N] */
  double* grad;     /* [15N]: (component*3 + axis)*N + i */
  double* gc;       /* [3N] gradients_centroid */
  double* extrap;   /* [5N] */
  double* cons;     /* [5N] */
  double* fluxes;   /* [5N] */
  double* volume;   /* [N] */
  double* loc;      /* [3N] */
  double* centroid; /* [3N] */
  double* v;        /* [3N] */
  double* v_rel;    /* [3N] */
  double* maor;     /* [N] max_a_over_r */
  double* msv;      /* [N] max_signal_velocity */
  double* dt;       /* [N] */
  int8_t* timebin;  /* [N] */
} ora_fields;
void ora_sim_get_fields(const ora_sim* s, ora_fields* f);
void ora_sim_set_fields(ora_sim* s, const ora_fields* f);
int64_t ora_sim_n_cells(const ora_sim* s);
int64_t ora_sim_n_faces(const ora_sim* s);
/* Copy the sim's current mesh (face arrays sized by n_faces, csr by 2*n_faces). Any pointer may be NULL. */
void ora_sim_get_mesh(const ora_sim* s, double* area, double* normal, double* centroid, int64_t* left,
                      int64_t* right, double* shift, int64_t* conn, int64_t* offset, int64_t* count);

/* Individual phases of the step, for per-phase parity tests.
 * fluxrec: [7F] per face: dt*F (5), v_max, a_over_r (FluxInfo without mflux). */
void ora_phase_predict(ora_sim* s, double dt);
void ora_phase_compute_fluxes(ora_sim* s, const uint8_t* active, double* fluxrec);
/* Status line sums (src/space.rs:1248-1280): mass+pending, energy+pending, plus momentum+pending. */
void ora_sim_conservation(const ora_sim* s, double* out5);
/* gradient_estimate/limit/apply (src/runner/hydrodynamics.rs:67-143) on the current primitives. */
void ora_phase_gradients(ora_sim* s, const uint8_t* active);
/* GravitySolver::compute_accelerations (src/gravity.rs:24-44, 64-69, 82-87, 116-148) on the current state:
 * out [3N] component-major.  The reference's step never calls it (optimal_order.rs:33), so a_grav stays
 * zero there; this returns the accelerations without storing them. */
void ora_gravity_accelerations(const ora_sim* s, double* out3n);
/* GravitySolver::compute_timesteps (src/gravity.rs:46-55, 71-74, 90-95, 107-110) with a_grav = 0: [N]. */
void ora_gravity_timesteps(const ora_sim* s, double* out);

/* Single functions of the path (ideal gas), so that tests can pin each against closed forms:
 * conserved -> primitive (physical_quantities.rs:213-226, part.rs:219-261), time_extrapolations
 * (part.rs:134-149), radius (part.rs:439-447), Particle::timestep (part.rs:42-84; in/out layout in
 * ora_runner.c), timestep_limit (part.rs:263-297) and the face velocity (flux.rs:75-78). */
int ora_t_c2p(double gamma, const double* cons5, double volume, double* prim5);
void ora_t_time_extrapolations(double gamma, const double* prim5, const double* grad15, const double* v_rel3,
                               double dt, double* out5);
double ora_t_radius(double volume, int dim);
int ora_t_particle_timestep(double gamma, double cfl, int motion, int dim, const double* in, double* out);
int8_t ora_t_timestep_limit(double dt, int8_t timebin, int8_t wakeup, uint64_t ti_current, double time_base,
                            double* dt_out);
void ora_t_face_velocity(const double* loc_l, const double* loc_r, const double* shift, const double* v_l,
                         const double* v_r, const double* fc, double* out3);

/* 1D tessellation (meshless_voronoi 1D semantics pinned by src/space.rs:1453-1491). Output arrays
 * sized: faces F = n+1 (non periodic) or n (periodic); conn 2n. */
int ora_voronoi_1d(int64_t n, const double* x, double box_x, int periodic, double* area, double* normal,
                   double* centroid, int64_t* left, int64_t* right, double* shift, int64_t* conn, int64_t* offset,
                   int64_t* count, double* volume, double* cell_centroid, double* search_radius,
                   int64_t* n_faces_out);

#ifdef __cplusplus
}
#endif
#endif
