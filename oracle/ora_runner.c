/* CPU oracle (test infrastructure only): particle operations, face flux exchange, runner phases,
 * integer timeline and engine loop.  Restates src/part.rs, src/flux.rs, src/gradients.rs,
 * src/runner/hydrodynamics.rs, src/runner/hydrodynamics/{optimal_order,godunov}.rs,
 * src/finite_volume_solver/euler_eqns_fvs.rs, src/timeline.rs, src/engine.rs and the parts of
 * src/space.rs on the path.  OpenMP loops sit where the reference uses rayon par_iter. */
#include <stdlib.h>
#include <string.h>
#include "ora_internal.h"

#define NUM_TIME_BINS 56
#define TIME_BIN_NEIGHBOUR_MAX_DELTA_BIN 2

/* ======================= timeline (src/timeline.rs) ======================= */
/* timeline.rs:18-24 */
uint64_t ora_get_integer_timestep(int8_t bin) {
  if (bin <= 0) return 0;
  return (uint64_t)1 << (bin + 1);
}
#define MAX_NR_TIMESTEPS ((uint64_t)1 << 57)
/* timeline.rs:31-33: (u64::BITS - leading_zeros - 2) in u32 arithmetic, then `as i8`. */
int8_t ora_get_time_bin(uint64_t dti) {
  uint32_t lz = dti ? (uint32_t)__builtin_clzll(dti) : 64u;
  uint32_t r = 64u - lz - 2u;
  return (int8_t)(uint8_t)(r & 0xFFu);
}
/* timeline.rs:53-67 */
uint64_t ora_get_integer_time_end(uint64_t ti_current, int8_t bin) {
  uint64_t dti = ora_get_integer_timestep(bin);
  if (dti == 0) return 0;
  uint64_t modulus = ti_current % dti;
  if (modulus == 0) return ti_current;
  return ti_current - modulus + dti;
}
/* timeline.rs:72-83 */
int8_t ora_get_max_active_bin(uint64_t ti) {
  if (ti == 0) return NUM_TIME_BINS;
  int bin = 1;
  while (bin < 62 && ((((uint64_t)1) << (bin + 1)) & ti) == 0) bin += 1;
  return (int8_t)bin;
}
/* Rust `f64 as u64`: saturating, NaN -> 0. */
uint64_t ora_saturating_f64_to_u64(double x) {
  if (!(x > 0.)) return 0; /* NaN, negative, zero */
  if (x >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)x;
}
/* timeline.rs:103-139 */
uint64_t ora_make_integer_timestep(double dt, int8_t old_bin, int8_t min_ngb_bin, uint64_t ti_current,
                                   double time_base_inv) {
  uint64_t new_dti = ora_saturating_f64_to_u64(dt * time_base_inv);
  int8_t b = ora_get_time_bin(new_dti);
  int8_t lim = (int8_t)(min_ngb_bin + TIME_BIN_NEIGHBOUR_MAX_DELTA_BIN);
  int8_t new_bin = b < lim ? b : lim;
  new_dti = ora_get_integer_timestep(new_bin);
  uint64_t current_dti = ora_get_integer_timestep(old_bin);
  if (old_bin > 0) {
    uint64_t t = 2 * current_dti;
    if (t < new_dti) new_dti = t;
  }
  uint64_t timeline_dti = MAX_NR_TIMESTEPS;
  while (new_dti < timeline_dti) timeline_dti /= 2;
  new_dti = timeline_dti;
  if (new_dti > current_dti) {
    uint64_t ti_end = ora_get_integer_time_end(ti_current, old_bin);
    if ((MAX_NR_TIMESTEPS - ti_end) % new_dti != 0) new_dti = current_dti;
  }
  return new_dti;
}
static double make_timestep(uint64_t dti, double time_base) { return (double)dti * time_base; }

/* ======================= particle (src/part.rs) ======================= */
typedef struct {
  st5 primitives;
  grad5 gradients;
  vec3 gradients_centroid;
  st5 extrapolations;
  st5 conserved;
  st5 fluxes;
  vec3 gravity_mflux;
  double volume;
  int64_t face_connections_offset, face_count;
  double search_radius;
  vec3 loc, centroid, v, v_rel;
  double max_a_over_r, max_signal_velocity;
  int8_t timebin;
  double dt, dvdt;
  vec3 a_grav;
} particle;

typedef struct {
  double area;
  vec3 normal, centroid, shift;
  int64_t left, right;
} face_t;

typedef struct {
  uint64_t ti_old, ti_current, ti_next;
  double time_base, time_base_inv;
} tsinfo;

struct ora_sim {
  ora_params p;
  gaslaw eos;
  int64_t n;
  particle* parts;
  int64_t nf;
  face_t* faces;
  int64_t nconn;
  int64_t* conn;
  /* engine (src/engine.rs:33-101) */
  tsinfo ts;
  double t_current;
  int64_t step_count;
  /* sync points queue */
  int64_t sp_cap, sp_head, sp_tail;
  uint64_t* sp_ti;
  int* sp_kind;
  uint64_t ti_snap, ti_between_snaps;
  int error;
  ora_counters cnt;
  int64_t last_n_active, last_n_face_active;
  ora_mesher mesher; /* 2D/3D moving meshes: the caller's tessellator */
  void* mesher_user;
  const uint8_t* mesh_mask;
};

enum { SP_DUMP = 0, SP_STEP = 3 };

static inline st5 grad_dot(const grad5* g, vec3 dx) {
  return st_make(v3_dot(g->g[0], dx), v3(v3_dot(g->g[1], dx), v3_dot(g->g[2], dx), v3_dot(g->g[3], dx)),
                 v3_dot(g->g[4], dx));
}
static inline int grad_finite(const grad5* g) {
  for (int i = 0; i < 5; ++i)
    if (!v3_finite(g->g[i])) return 0;
  return 1;
}
static inline double div_v(const grad5* g) { return g->g[1].x + g->g[2].y + g->g[3].z; }
static inline vec3 curl_v(const grad5* g) {
  return v3(g->g[3].y - g->g[2].z, g->g[1].z - g->g[3].x, g->g[2].x - g->g[1].y);
}

/* part.rs:134-149 */
static st5 time_extrapolations(const particle* p, double dt, const gaslaw* eos) {
  double rho = p->primitives.c[0];
  if (rho > 0.) {
    double rho_inv = 1. / rho;
    double P = p->primitives.c[4];
    double dv = div_v(&p->gradients);
    st5 s = st_make(rho * dv + v3_dot(p->v_rel, p->gradients.g[0]),
                    v3_add(v3_scale(dv, p->v_rel), v3_scale(rho_inv, p->gradients.g[4])),
                    eos->gamma.gamma * P * dv + v3_dot(p->v_rel, p->gradients.g[4]));
    return st_scale(-dt, s);
  }
  return st_vacuum();
}
/* part.rs:123-132 */
static void extrapolate_state(particle* p, double dt, const gaslaw* eos, ora_counters* cnt) {
  p->extrapolations = st_add(p->extrapolations, time_extrapolations(p, dt, eos));
  ora_check_physical(&p->primitives, cnt);
  double volume = p->volume + p->dvdt * dt;
  double lo = 0.5 * p->volume, hi = 2. * p->volume;
  /* f64::clamp: NaN stays NaN */
  if (volume < lo) volume = lo;
  if (volume > hi) volume = hi;
  p->volume = volume;
}
/* part.rs:439-447 */
static double radius(const particle* p, int dim) {
  if (dim == 1) return 0.5 * p->volume;
  if (dim == 2) return sqrt(0.31830988618379067154 * p->volume);
  return pow(0.25 * 3. * 0.31830988618379067154 * p->volume, 1. / 3.);
}
/* part.rs:314-317 */
static double particle_internal_energy(const particle* p) {
  return (p->conserved.c[4] - 0.5 * v3_dot(st_vec(p->conserved), st_vec(p->primitives))) / p->conserved.c[0];
}
/* part.rs:328-372 */
static int set_particle_velocity(particle* p, vec3 fluid_v, double sound_speed, int motion, int dim) {
  switch (motion) {
    case ORA_FIXED:
      p->v = V3_ZERO;
      break;
    case ORA_FLUID:
      p->v = fluid_v;
      break;
    case ORA_STEER: {
      vec3 d = v3_sub(p->centroid, p->loc);
      double abs_d = v3_len(d);
      double eta = 0.25;
      double eta_r = eta * radius(p, dim);
      double xi = 0.99;
      if (abs_d > 0.9 * eta_r) {
        double fac = xi * sound_speed / abs_d;
        if (abs_d < 1.1 * eta_r) fac *= 5. * (abs_d - 0.9 * eta_r) / eta_r;
        p->v = v3_add(fluid_v, v3_scale(fac, d));
      } else {
        p->v = fluid_v;
      }
      break;
    }
    case ORA_STEER_PAKMOR: {
      double alpha = 2. / (double)dim * p->max_a_over_r;
      double beta = 2.25;
      double eta = rmax(0., rmin(0.5, 2. / beta * (alpha - 0.75 * beta)));
      vec3 dd = v3_sub(p->centroid, p->loc);
      double lr = 1. / v3_len(dd);
      vec3 d = v3_scale(lr, dd);
      double v_reg = radius(p, dim) * v3_len(curl_v(&p->gradients));
      vec3 vr = v3_scale(eta * rmax(sound_speed, v_reg), d);
      p->v = v3_add(fluid_v, vr);
      break;
    }
  }
  if (!v3_finite(p->v)) return ORA_ERR_NONFINITE;
  p->v_rel = v3_sub(fluid_v, p->v);
  p->max_a_over_r = 0.;
  return ORA_OK;
}
/* part.rs:42-84 */
static double particle_timestep(particle* p, double cfl, int motion, const gaslaw* eos, int dim, int* err) {
  if (p->conserved.c[0] == 0.) {
    p->v = V3_ZERO;
    return INFINITY;
  }
  double m_inv = 1. / p->conserved.c[0];
  st5 dvs = grad_dot(&p->gradients, v3_sub(p->loc, p->gradients_centroid));
  vec3 d_v = st_vec(dvs);
  vec3 fluid_v = v3_add(v3_scale(m_inv, st_vec(p->conserved)), d_v);
  double sound_speed = gas_sound_speed(
      eos, gas_pressure_from_internal_energy(eos, particle_internal_energy(p), p->primitives.c[0]),
      p->volume * m_inv);
  int e = set_particle_velocity(p, fluid_v, sound_speed, motion, dim);
  if (e) *err = e;
  double v_max = rmax(p->max_signal_velocity, v3_len(p->v_rel) + sound_speed);
  p->max_signal_velocity = 0.;
  if (v_max > 0.) return cfl * radius(p, dim) / v_max;
  return INFINITY;
}
/* part.rs:303-312 */
static void grav_kick(particle* p) {
  double mass = p->conserved.c[0];
  vec3 momentum = st_vec(p->conserved);
  vec3 gkf = v3_scale(0.5 * p->dt, p->a_grav);
  p->conserved = st_add(p->conserved, st_make(0., v3_scale(mass, gkf), v3_dot(gkf, v3_sub(momentum, p->gravity_mflux))));
}
/* part.rs:87-104 */
static void hydro_kick1(particle* p, int motion, int dim) {
  if (motion == ORA_FIXED) return;
  double rho = p->primitives.c[0];
  if (rho > 0.) {
    double kick_fac = 0.5 * p->dt / rho;
    for (int i = 0; i < dim; ++i) v3_set(&p->v, i, v3_get(p->v, i) - kick_fac * v3_get(p->gradients.g[4], i));
  }
}
/* part.rs:178-198 */
static int apply_flux(particle* p, ora_counters* cnt) {
  p->conserved = st_add(p->conserved, p->fluxes);
  for (int i = 0; i < 5; ++i)
    if (!isfinite(p->conserved.c[i])) return ORA_ERR_NONFINITE;
  if (p->conserved.c[0] < 0.) {
    cnt->apply_flux_resets++;
    p->conserved = st_vacuum();
  }
  if (p->conserved.c[4] < 0.) {
    cnt->apply_flux_resets++;
    p->conserved = st_vacuum();
  }
  p->fluxes = st_vacuum();
  p->gravity_mflux = V3_ZERO;
  return ORA_OK;
}
/* physical_quantities.rs:213-226 + part.rs:219-261 */
static int convert_conserved_to_primitive(particle* p, const gaslaw* eos) {
  st5 c = p->conserved;
  if (!(c.c[0] >= 0.)) return ORA_ERR_NEG_MASS;
  if (!(c.c[4] >= 0.)) return ORA_ERR_NEG_ENERGY;
  if (c.c[0] == 0.) {
    if (c.c[4] != 0.) return ORA_ERR_NEG_ENERGY;
    p->primitives = st_vacuum();
  } else if (c.c[0] > 0.) {
    double m_inv = 1. / c.c[0];
    double density = c.c[0] / p->volume;
    vec3 velocity = v3_scale(m_inv, st_vec(c));
    double thermal = c.c[4] - 0.5 * v3_len2(st_vec(c)) * m_inv;
    double u = thermal * m_inv;
    double pressure = gas_pressure_from_internal_energy(eos, u, density);
    if (!(density >= 0.)) return ORA_ERR_NEG_PRESSURE;
    if (!(pressure >= 0.)) return ORA_ERR_NEG_PRESSURE;
    p->primitives = st_make(density, velocity, pressure);
  } else {
    p->primitives = st_vacuum();
  }
  if (!(p->primitives.c[0] >= 0.) || !(p->primitives.c[4] >= 0.)) return ORA_ERR_NEG_PRESSURE;
  for (int i = 0; i < 5; ++i)
    if (!isfinite(p->primitives.c[i])) return ORA_ERR_NONFINITE;
  return ORA_OK;
}
/* part.rs:374-388 */
static particle particle_reflect(const particle* p, vec3 around, vec3 n) {
  particle r = *p;
  r.loc = v3_add(r.loc, v3_scale(2. * v3_dot(v3_sub(around, p->loc), n), n));
  r.centroid = v3_add(r.centroid, v3_scale(2. * v3_dot(v3_sub(around, p->centroid), n), n));
  r.gradients_centroid =
      v3_add(r.gradients_centroid, v3_scale(2. * v3_dot(v3_sub(around, p->gradients_centroid), n), n));
  r.v = v3_sub(r.v, v3_scale(2. * v3_dot(r.v, n), n));
  return r;
}
/* physical_quantities.rs:239-242 */
static st5 prim_reflect(st5 s, vec3 n) {
  vec3 v = st_vec(s);
  vec3 rv = v3_sub(v, v3_scale(2. * v3_dot(v, n), n));
  return st_make(s.c[0], rv, s.c[4]);
}
/* part.rs:390-429 */
static void reflect_gradients(particle* p, vec3 n) {
  grad5 g = p->gradients;
  g.g[0] = v3_sub(g.g[0], v3_scale(2. * v3_dot(g.g[0], n), n));
  g.g[4] = v3_sub(g.g[4], v3_scale(2. * v3_dot(g.g[0], n), n));
  vec3 dv_dx = v3(g.g[1].x, g.g[2].x, g.g[3].x);
  vec3 dv_dy = v3(g.g[1].y, g.g[2].y, g.g[3].y);
  vec3 dv_dz = v3(g.g[1].z, g.g[2].z, g.g[3].z);
  vec3 dv_dot_n = v3(v3_dot(dv_dx, n), v3_dot(dv_dy, n), v3_dot(dv_dz, n));
  g.g[1] = v3_sub(g.g[1], v3_scale(n.x, v3_scale(2., dv_dot_n)));
  g.g[2] = v3_sub(g.g[2], v3_scale(n.y, v3_scale(2., dv_dot_n)));
  g.g[3] = v3_sub(g.g[3], v3_scale(n.z, v3_scale(2., dv_dot_n)));
  for (int k = 1; k <= 3; ++k) g.g[k] = v3_sub(g.g[k], v3_scale(2. * v3_dot(g.g[k], n), n));
  p->gradients = g;
}

/* space.rs:213-229 */
static particle get_boundary_part(const ora_sim* s, const particle* p, const face_t* f) {
  particle r = particle_reflect(p, f->centroid, f->normal);
  if (s->p.boundary == ORA_REFLECTIVE) {
    r.primitives = prim_reflect(r.primitives, f->normal);
  } else if (s->p.boundary == ORA_VACUUM) {
    r.primitives = st_vacuum();
  }
  return r;
}

/* ======================= flux (src/flux.rs) ======================= */
typedef struct {
  st5 fluxes;
  vec3 mflux;
  double v_max, a_over_r;
} flux_info;

static flux_info flux_zero(void) {
  flux_info f;
  f.fluxes = st_vacuum();
  f.mflux = V3_ZERO;
  f.v_max = 0.;
  f.a_over_r = 0.;
  return f;
}

/* Face velocity (flux.rs:75-78): the mean generator velocity minus the part of their relative velocity
 * that moves the face point along dx. */
static vec3 face_velocity(vec3 loc_l, vec3 loc_r, vec3 shift, vec3 v_l, vec3 v_r, vec3 fc) {
  vec3 dx = v3_sub(v3_add(loc_r, shift), loc_l);
  vec3 midpoint = v3_scale(0.5, v3_add(v3_add(loc_l, loc_r), shift));
  double fac = v3_dot(v3_sub(v_r, v_l), v3_sub(fc, midpoint)) / v3_len2(dx);
  return v3_sub(v3_scale(0.5, v3_add(v_l, v_r)), v3_scale(fac, dx));
}

/* flux.rs:23-100 */
static int flux_exchange(const particle* left, const particle* right, double dt, const face_t* face,
                         double time_extrapolate_fac, const gaslaw* eos, const ora_riemann* rs, flux_info* out,
                         ora_counters* cnt) {
  vec3 dx_left = v3_sub(face->centroid, left->gradients_centroid);
  vec3 shift = face->shift;
  vec3 dx_right = v3_sub(v3_sub(face->centroid, right->gradients_centroid), shift);
  double dx_gradients = v3_len(v3_sub(dx_left, dx_right));
  vec3 dx_centroid = v3_sub(v3_add(right->centroid, shift), left->centroid);
  vec3 dx = v3_sub(v3_add(right->loc, shift), left->loc);

  double v_max = 0.;
  if (left->primitives.c[0] > 0.) v_max += gas_sound_speed(eos, left->primitives.c[4], 1. / left->primitives.c[0]);
  if (right->primitives.c[0] > 0.)
    v_max += gas_sound_speed(eos, right->primitives.c[4], 1. / right->primitives.c[0]);
  v_max -= rmin(v3_dot(v3_sub(st_vec(right->primitives), st_vec(left->primitives)), dx_centroid), 0.);

  double dt_extrapolate = -time_extrapolate_fac * dt;
  st5 left_dash = st_add(st_add(st_add(left->primitives, grad_dot(&left->gradients, dx_left)), left->extrapolations),
                         time_extrapolations(left, dt_extrapolate, eos));
  st5 right_dash =
      st_add(st_add(st_add(right->primitives, grad_dot(&right->gradients, dx_right)), right->extrapolations),
             time_extrapolations(right, dt_extrapolate, eos));
  st5 prim_left = ora_pairwise(left->primitives, right->primitives, left_dash, v3_len(dx_left) / dx_gradients, cnt);
  st5 prim_right =
      ora_pairwise(right->primitives, left->primitives, right_dash, v3_len(dx_right) / dx_gradients, cnt);

  vec3 v_face = face_velocity(left->loc, right->loc, shift, left->v, right->v, face->centroid);

  /* State::boost (physical_quantities.rs:228-236) */
  vec3 nvf = v3_neg(v_face);
  st5 bl = prim_left, br = prim_right;
  if (bl.c[0] > 0.) bl = st_make(bl.c[0], v3_add(st_vec(bl), nvf), bl.c[4]);
  if (br.c[0] > 0.) br = st_make(br.c[0], v3_add(st_vec(br), nvf), br.c[4]);
  st5 F;
  int e = ora_riemann_solve(rs, bl, br, v_face, face->normal, eos, &F, cnt);
  if (e) return e;
  st5 fluxes = st_scale(face->area, F);
  out->fluxes = st_scale(dt, fluxes);
  out->mflux = v3_scale(fluxes.c[0], dx);
  out->v_max = v_max;
  out->a_over_r = face->area / v3_len(dx);
  return ORA_OK;
}

/* flux.rs:102-195 */
static int flux_exchange_boundary(const ora_sim* s, const particle* part, const face_t* face,
                                  double time_extrapolate_fac, flux_info* out, ora_counters* cnt) {
  const gaslaw* eos = &s->eos;
  particle reflected = particle_reflect(part, face->centroid, face->normal);
  vec3 dx_face = v3_sub(face->centroid, part->gradients_centroid);
  double dx_gradients = v3_len(v3_sub(reflected.gradients_centroid, part->gradients_centroid));
  vec3 dx_centroid = v3_sub(reflected.centroid, part->centroid);
  vec3 dx = v3_sub(reflected.loc, part->loc);
  double v_max = 0.;
  if (part->primitives.c[0] > 0.) v_max += gas_sound_speed(eos, part->primitives.c[4], 1. / part->primitives.c[0]);
  double dt_extrapolate = -time_extrapolate_fac * part->dt;
  st5 dash = st_add(st_add(st_add(part->primitives, grad_dot(&part->gradients, dx_face)), part->extrapolations),
                    time_extrapolations(part, dt_extrapolate, eos));
  st5 boundary_state;
  double dx_fac = v3_len(dx_face) / dx_gradients;
  switch (s->p.boundary) {
    case ORA_REFLECTIVE:
      reflected.primitives = prim_reflect(reflected.primitives, face->normal);
      dash = ora_pairwise(part->primitives, reflected.primitives, dash, dx_fac, cnt);
      v_max -= rmin(v3_dot(v3_sub(st_vec(reflected.primitives), st_vec(part->primitives)), dx_centroid), 0.);
      boundary_state = prim_reflect(dash, face->normal);
      break;
    case ORA_OPEN:
      dash = ora_pairwise(part->primitives, part->primitives, dash, dx_fac, cnt);
      boundary_state = dash;
      break;
    case ORA_VACUUM: {
      st5 vac = st_vacuum();
      dash = ora_pairwise(part->primitives, vac, dash, dx_fac, cnt);
      v_max += rmin(v3_dot(st_vec(part->primitives), dx_centroid), 0.);
      boundary_state = vac;
      break;
    }
    default:
      return ORA_ERR_UNSUPPORTED;
  }
  st5 F;
  int e = ora_riemann_solve(&s->p.riemann, dash, boundary_state, V3_ZERO, face->normal, eos, &F, cnt);
  if (e) return e;
  st5 fluxes = st_scale(face->area, F);
  out->fluxes = st_scale(part->dt, fluxes);
  out->mflux = v3_scale(fluxes.c[0], dx);
  out->v_max = v_max;
  out->a_over_r = face->area / v3_len(dx);
  return ORA_OK;
}

/* ======================= runner phases ======================= */
static void merge_counters(ora_counters* dst, const ora_counters* src) {
  dst->check_physical_resets += src->check_physical_resets;
  dst->apply_flux_resets += src->apply_flux_resets;
  dst->gradient_fallbacks += src->gradient_fallbacks;
  dst->brent_calls += src->brent_calls;
}
static void set_error(ora_sim* s, int e) {
  if (e && !s->error) s->error = e;
}

/* TimestepInfo (src/engine.rs:33-81) */
static int bin_is_ending(const tsinfo* t, int8_t bin) { return bin <= ora_get_max_active_bin(t->ti_current); }
static double dt_from_dti(const tsinfo* t, uint64_t dti) { return make_timestep(dti, t->time_base); }

/* EulerEqnsFvs::predict (euler_eqns_fvs.rs:20-24) */
void ora_phase_predict(ora_sim* s, double dt) {
  ora_counters total = {0, 0, 0, 0};
#pragma omp parallel
  {
    ora_counters c = {0, 0, 0, 0};
#pragma omp for schedule(static)
    for (int64_t i = 0; i < s->n; ++i) extrapolate_state(&s->parts[i], dt, &s->eos, &c);
#pragma omp critical
    merge_counters(&total, &c);
  }
  merge_counters(&s->cnt, &total);
}

/* EulerEqnsFvs::compute_fluxes (euler_eqns_fvs.rs:26-48) */
static void compute_fluxes(ora_sim* s, const uint8_t* active, flux_info* out) {
  ora_counters total = {0, 0, 0, 0};
  int err = 0;
  int64_t nfa = 0;
#pragma omp parallel reduction(+ : nfa)
  {
    ora_counters c = {0, 0, 0, 0};
    int lerr = 0;
#pragma omp for schedule(static)
    for (int64_t fi = 0; fi < s->nf; ++fi) {
      const face_t* face = &s->faces[fi];
      const particle* left = &s->parts[face->left];
      int la = active[face->left];
      int e = 0;
      if (face->right >= 0) {
        const particle* right = &s->parts[face->right];
        int ra = active[face->right];
        if ((!la && !ra) || (right->dt < left->dt && !ra) || (left->dt < right->dt && !la)) {
          out[fi] = flux_zero();
          continue;
        }
        double dt = rmin(left->dt, right->dt);
        e = flux_exchange(left, right, dt, face, 0.5, &s->eos, &s->p.riemann, &out[fi], &c);
        nfa++;
      } else if (la) {
        e = flux_exchange_boundary(s, left, face, 0.5, &out[fi], &c);
        nfa++;
      } else {
        out[fi] = flux_zero();
      }
      if (e && !lerr) lerr = e;
    }
#pragma omp critical
    {
      merge_counters(&total, &c);
      if (lerr && !err) err = lerr;
    }
  }
  merge_counters(&s->cnt, &total);
  set_error(s, err);
  s->last_n_face_active = nfa;
}

void ora_phase_compute_fluxes(ora_sim* s, const uint8_t* active, double* fluxrec) {
  flux_info* fl = (flux_info*)malloc(sizeof(flux_info) * (size_t)(s->nf ? s->nf : 1));
  compute_fluxes(s, active, fl);
  for (int64_t f = 0; f < s->nf; ++f) {
    for (int k = 0; k < 5; ++k) fluxrec[7 * f + k] = fl[f].fluxes.c[k];
    fluxrec[7 * f + 5] = fl[f].v_max;
    fluxrec[7 * f + 6] = fl[f].a_over_r;
  }
  free(fl);
}

/* part.rs:160-176 */
static void update_fluxes_left(particle* p, const flux_info* fi, int active) {
  p->fluxes = st_sub(p->fluxes, fi->fluxes);
  p->gravity_mflux = v3_sub(p->gravity_mflux, fi->mflux);
  if (active) {
    p->max_signal_velocity = rmax(fi->v_max, p->max_signal_velocity);
    p->max_a_over_r = rmax(p->max_a_over_r, fi->a_over_r);
  }
}
static void update_fluxes_right(particle* p, const flux_info* fi, int active) {
  p->fluxes = st_add(p->fluxes, fi->fluxes);
  p->gravity_mflux = v3_sub(p->gravity_mflux, fi->mflux);
  if (active) {
    p->max_signal_velocity = rmax(fi->v_max, p->max_signal_velocity);
    p->max_a_over_r = rmax(p->max_a_over_r, fi->a_over_r);
  }
}

/* runner/hydrodynamics.rs:13-39 */
static void apply_fluxes(ora_sim* s, const flux_info* fluxes, const uint8_t* active) {
  ora_counters total = {0, 0, 0, 0};
  int err = 0;
#pragma omp parallel
  {
    ora_counters c = {0, 0, 0, 0};
    int lerr = 0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < s->n; ++i) {
      particle* p = &s->parts[i];
      for (int64_t k = 0; k < p->face_count; ++k) {
        int64_t idx = s->conn[p->face_connections_offset + k];
        const face_t* f = &s->faces[idx];
        if (f->left == i)
          update_fluxes_left(p, &fluxes[idx], active[i]);
        else
          update_fluxes_right(p, &fluxes[idx], active[i]);
      }
      if (active[i]) {
        int e = apply_flux(p, &c);
        if (e && !lerr) lerr = e;
      }
    }
#pragma omp critical
    {
      merge_counters(&total, &c);
      if (lerr && !err) err = lerr;
    }
  }
  merge_counters(&s->cnt, &total);
  set_error(s, err);
}

static void kick2(ora_sim* s, const uint8_t* active) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; ++i)
    if (active[i]) grav_kick(&s->parts[i]);
}

static void kick1(ora_sim* s, const uint8_t* active) {
  int dim = s->p.dim;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; ++i)
    if (active[i]) {
      hydro_kick1(&s->parts[i], s->p.motion, dim);
      grav_kick(&s->parts[i]);
    }
}

/* euler_eqns_fvs.rs:63-69 */
static void c2p(ora_sim* s, const uint8_t* active) {
  int err = 0;
#pragma omp parallel
  {
    int lerr = 0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < s->n; ++i)
      if (active[i]) {
        int e = convert_conserved_to_primitive(&s->parts[i], &s->eos);
        if (e && !lerr) lerr = e;
      }
#pragma omp critical
    if (lerr && !err) err = lerr;
  }
  set_error(s, err);
}

/* gradients.rs:8-56 (GradientData) with glam 0.27 DMat3 semantics: columns x_axis, y_axis, z_axis */
typedef struct { vec3 x, y, z; } mat3;

static void gradient_data_init(int dim, grad5* g, mat3* m) {
  for (int i = 0; i < 5; ++i) g->g[i] = V3_ZERO;
  m->x = V3_ZERO;
  m->y = V3_ZERO;
  m->z = V3_ZERO;
  if (dim == 1) { /* IDENTITY with x_axis.x = 0 */
    m->y.y = 1.;
    m->z.z = 1.;
  } else if (dim == 2) {
    m->z.z = 1.;
  }
}
static void gradient_data_collect(grad5* g, mat3* m, st5 sl, st5 sr, double w, vec3 ds) {
  for (int i = 0; i < 5; ++i) g->g[i] = v3_add(g->g[i], v3_scale(w * (sr.c[i] - sl.c[i]), ds));
  mat3 add;
  add.x = v3_scale(w, v3_scale(ds.x, ds));
  add.y = v3_scale(w, v3_scale(ds.y, ds));
  add.z = v3_scale(w, v3_scale(ds.z, ds));
  m->x = v3_add(m->x, add.x);
  m->y = v3_add(m->y, add.y);
  m->z = v3_add(m->z, add.z);
}
/* glam DMat3::inverse: cofactor columns via cross products, det = z_axis . (x_axis x y_axis),
 * multiply by det.recip(), transpose.  Parity unpinned (restated from glam's public source). */
static mat3 mat3_inverse(mat3 a) {
  vec3 t0 = v3_cross(a.y, a.z);
  vec3 t1 = v3_cross(a.z, a.x);
  vec3 t2 = v3_cross(a.x, a.y);
  double det = v3_dot(a.z, t2);
  double inv_det = 1. / det;
  vec3 c0 = v3(t0.x * inv_det, t0.y * inv_det, t0.z * inv_det);
  vec3 c1 = v3(t1.x * inv_det, t1.y * inv_det, t1.z * inv_det);
  vec3 c2 = v3(t2.x * inv_det, t2.y * inv_det, t2.z * inv_det);
  mat3 r; /* transpose of from_cols(c0, c1, c2) */
  r.x = v3(c0.x, c1.x, c2.x);
  r.y = v3(c0.y, c1.y, c2.y);
  r.z = v3(c0.z, c1.z, c2.z);
  return r;
}
/* glam DMat3::mul_vec3: (x_axis*v.x + y_axis*v.y) + z_axis*v.z */
static vec3 mat3_mul_vec3(mat3 m, vec3 v) {
  vec3 r = v3_scale(v.x, m.x);
  r = v3_add(r, v3_scale(v.y, m.y));
  r = v3_add(r, v3_scale(v.z, m.z));
  return r;
}
static grad5 gradient_data_finalize(grad5 g, mat3 m, ora_counters* cnt) {
  mat3 inv = mat3_inverse(m);
  for (int i = 0; i < 5; ++i) g.g[i] = mat3_mul_vec3(inv, g.g[i]);
  if (!grad_finite(&g)) {
    cnt->gradient_fallbacks++;
    for (int i = 0; i < 5; ++i) g.g[i] = V3_ZERO;
  }
  return g;
}

/* gradients.rs:58-103 (LimiterData) */
static double limit_single_quantity(double d_max, double d_min, double e_max, double e_min) {
  if (e_max == 0. || e_min == 0.) return 1.;
  return rmin(1.0, rmin(d_max / e_max, d_min / e_min));
}

/* runner/hydrodynamics.rs:67-143: estimate, limit, apply (computed into a buffer first). */
static void gradients_phase(ora_sim* s, const uint8_t* active) {
  int64_t n = s->n;
  grad5* out = (grad5*)malloc(sizeof(grad5) * (size_t)(n ? n : 1));
  ora_counters total = {0, 0, 0, 0};
#pragma omp parallel
  {
    ora_counters c = {0, 0, 0, 0};
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      if (!active[i]) continue;
      const particle* part = &s->parts[i];
      vec3 centroid = part->centroid;
      grad5 g;
      mat3 m;
      gradient_data_init(s->p.dim, &g, &m);
      for (int64_t k = 0; k < part->face_count; ++k) {
        const face_t* face = &s->faces[s->conn[part->face_connections_offset + k]];
        double area = face->area;
        vec3 shift = (i == face->left) ? face->shift : v3_neg(face->shift);
        int64_t other_idx = (i == face->left) ? face->right : face->left;
        if (other_idx >= 0) {
          const particle* other = &s->parts[other_idx];
          vec3 ds = v3_sub(v3_add(other->centroid, shift), centroid);
          gradient_data_collect(&g, &m, part->primitives, other->primitives, area / v3_len2(ds), ds);
        } else {
          particle other = get_boundary_part(s, part, face);
          vec3 ds = v3_sub(v3_add(other.centroid, shift), centroid);
          gradient_data_collect(&g, &m, part->primitives, other.primitives, area / v3_len2(ds), ds);
        }
      }
      out[i] = gradient_data_finalize(g, m, &c);
    }
#pragma omp critical
    merge_counters(&total, &c);
  }
  merge_counters(&s->cnt, &total);

  /* gradient_limit (hydrodynamics.rs:104-134) */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (!active[i]) continue;
    const particle* part = &s->parts[i];
    vec3 centroid = part->centroid;
    st5 lmin = part->primitives, lmax = part->primitives;
    st5 emin, emax;
    for (int q = 0; q < 5; ++q) {
      emin.c[q] = INFINITY;
      emax.c[q] = -INFINITY;
    }
    for (int64_t k = 0; k < part->face_count; ++k) {
      const face_t* face = &s->faces[s->conn[part->face_connections_offset + k]];
      vec3 shift = (i == face->left) ? face->shift : v3_neg(face->shift);
      st5 extrapolated = grad_dot(&out[i], v3_sub(v3_sub(face->centroid, centroid), shift));
      int64_t other_idx = (i == face->left) ? face->right : face->left;
      st5 other_prim;
      if (other_idx >= 0)
        other_prim = s->parts[other_idx].primitives;
      else
        other_prim = get_boundary_part(s, part, face).primitives;
      for (int q = 0; q < 5; ++q) {
        lmin.c[q] = rmin(lmin.c[q], other_prim.c[q]);
        lmax.c[q] = rmax(lmax.c[q], other_prim.c[q]);
        emin.c[q] = rmin(emin.c[q], extrapolated.c[q]);
        emax.c[q] = rmax(emax.c[q], extrapolated.c[q]);
      }
    }
    for (int q = 0; q < 5; ++q) {
      double alpha = limit_single_quantity(lmax.c[q] - part->primitives.c[q], lmin.c[q] - part->primitives.c[q],
                                           emax.c[q], emin.c[q]);
      out[i].g[q] = v3_scale(alpha, out[i].g[q]);
    }
  }
  /* gradient_apply (hydrodynamics.rs:136-143) */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (!active[i]) continue;
    s->parts[i].gradients = out[i];
    s->parts[i].gradients_centroid = s->parts[i].centroid;
  }
  free(out);
}

/* part.rs:263-297 */
static void timestep_limit(particle* p, int8_t wakeup, const tsinfo* t) {
  if (wakeup >= p->timebin) return;
  if (p->timebin <= ora_get_max_active_bin(t->ti_current)) {
    p->timebin = wakeup;
    p->dt = dt_from_dti(t, ora_get_integer_timestep(p->timebin));
  } else {
    uint64_t ti_end_old = ora_get_integer_time_end(t->ti_current, p->timebin);
    uint64_t ti_current = t->ti_current;
    p->dt -= dt_from_dti(t, ti_end_old - ti_current);
    p->timebin = wakeup;
    uint64_t dti_new = ora_get_integer_timestep(p->timebin);
    uint64_t ti_end = ora_get_integer_time_end(t->ti_current, p->timebin);
    p->dt += (ti_end == ti_current) ? dt_from_dti(t, dti_new) : dt_from_dti(t, ti_end - ti_current);
  }
}

/* ======================= gravity (src/gravity.rs) ======================= */
/* KeplerianPotential::acceleration (gravity.rs:127-138) */
static vec3 kepler_acceleration(const ora_params* p, vec3 pos) {
  vec3 dx = v3_sub(pos, v3(p->grav_position[0], p->grav_position[1], p->grav_position[2]));
  double r2 = v3_len2(dx);
  double r = sqrt(r2);
  if (r < p->grav_softening) r2 = r2 + p->grav_softening * p->grav_softening;
  double d = r * r2;
  return v3(-dx.x / d, -dx.y / d, -dx.z / d);
}
/* KeplerianPotential::get_timestep + circular_velocity (gravity.rs:121-125, 140-148), eta = 0.5 */
static double kepler_timestep(const ora_params* p, const particle* part) {
  vec3 dx = v3_sub(part->loc, v3(p->grav_position[0], p->grav_position[1], p->grav_position[2]));
  double r = v3_len(dx);
  double r_softened = r < p->grav_softening ? r + p->grav_softening : r;
  return 0.5 * radius(part, 2) / sqrt(1. / r_softened);
}
/* a particle's gravity timestep with a_grav = 0 (SelfGravity and Constant give infinity) */
static double gravity_timestep(const ora_params* p, const particle* part) {
  if (p->gravity == ORA_GRAV_KEPLERIAN) return kepler_timestep(p, part);
  return INFINITY;
}

void ora_gravity_accelerations(const ora_sim* s, double* out3n) {
  int64_t n = s->n;
  const ora_params* p = &s->p;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    vec3 a = V3_ZERO;
    const particle* pi = &s->parts[i];
    if (p->gravity == ORA_GRAV_SELF) {
      /* SelfGravity::compute_accelerations (gravity.rs:24-44): only j > i contribute */
      for (int64_t j = i + 1; j < n; ++j) {
        const particle* pj = &s->parts[j];
        double r = v3_len(v3_sub(pj->loc, pi->loc)) + p->grav_softening;
        vec3 dir = v3_sub(pj->loc, pi->loc);
        vec3 md = v3_scale(pj->conserved.c[0], dir);
        double r3 = r * r * r;
        a = v3_add(a, v3(md.x / r3, md.y / r3, md.z / r3));
      }
    } else if (p->gravity == ORA_GRAV_CONSTANT) {
      a = v3(p->grav_acceleration[0], p->grav_acceleration[1], p->grav_acceleration[2]);
    } else if (p->gravity == ORA_GRAV_KEPLERIAN) {
      a = kepler_acceleration(p, pi->loc);
    }
    out3n[i] = a.x;
    out3n[n + i] = a.y;
    out3n[2 * n + i] = a.z;
  }
}

void ora_gravity_timesteps(const ora_sim* s, double* out) {
  for (int64_t i = 0; i < s->n; ++i) out[i] = gravity_timestep(&s->p, &s->parts[i]);
}

/* OptimalOrderRunner::step (optimal_order.rs:15-64) / GodunovHydroRunner::step (godunov.rs:14-58),
 * without the tessellation (the caller rebuilds the mesh between predict and fluxes). */
static int mesh_update_1d(ora_sim* s, const uint8_t* mask);
static int mesh_update_caller(ora_sim* s, const uint8_t* mask);

int ora_sim_runner_step(ora_sim* s, uint64_t* ti_next_out) {
  int64_t n = s->n;
  const tsinfo* t = &s->ts;
  uint8_t* active = (uint8_t*)malloc((size_t)(n ? n : 1));
  int64_t nact = 0;
  for (int64_t i = 0; i < n; ++i) {
    active[i] = (uint8_t)bin_is_ending(t, s->parts[i].timebin); /* optimal_order.rs:17 */
    nact += active[i];
  }
  s->last_n_active = nact;
  double dt = dt_from_dti(t, t->ti_current - t->ti_old);
  ora_phase_predict(s, dt);
  if (s->p.motion != ORA_FIXED) {
    int e = s->p.dim == 1 ? mesh_update_1d(s, active) : mesh_update_caller(s, active);
    if (e) set_error(s, e);
  }
  flux_info* fl = (flux_info*)malloc(sizeof(flux_info) * (size_t)(s->nf ? s->nf : 1));
  compute_fluxes(s, active, fl);
  apply_fluxes(s, fl, active);
  free(fl);
  kick2(s, active);
  c2p(s, active);
  if (s->p.runner == ORA_OPTIMAL_ORDER) gradients_phase(s, active);

  /* compute_timesteps (euler_eqns_fvs.rs:50-61) + timesteps_apply (hydrodynamics.rs:145-158) */
  uint64_t min_dti = UINT64_MAX;
  int err = 0;
#pragma omp parallel
  {
    uint64_t lmin = UINT64_MAX;
    int lerr = 0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      if (!active[i]) continue;
      particle* p = &s->parts[i];
      double dts = particle_timestep(p, s->p.cfl, s->p.motion, &s->eos, s->p.dim, &lerr);
      /* the minimum with the gravity solver's timestep (optimal_order.rs:43-47; f64::min) */
      if (s->p.gravity != ORA_GRAV_NONE) dts = rmin(dts, gravity_timestep(&s->p, p));
      uint64_t dti = ora_make_integer_timestep(dts, p->timebin, NUM_TIME_BINS, t->ti_current, t->time_base_inv);
      p->dt = make_timestep(dti, t->time_base);
      p->timebin = ora_get_time_bin(dti);
      if (dti < lmin) lmin = dti;
    }
#pragma omp critical
    {
      if (lmin < min_dti) min_dti = lmin;
      if (lerr && !err) err = lerr;
    }
  }
  set_error(s, err);
  if (nact == 0) min_dti = MAX_NR_TIMESTEPS; /* the reference panics: "At least one particle must be active" */
  uint64_t ti_next = t->ti_current + min_dti;

  if (s->p.sync_all) {
    /* timestep_sync (hydrodynamics.rs:186-191) */
    int8_t minb = 127;
    for (int64_t i = 0; i < n; ++i)
      if (s->parts[i].timebin < minb) minb = s->parts[i].timebin;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) timestep_limit(&s->parts[i], minb, t);
  } else {
    /* timestep_limiter (hydrodynamics.rs:160-184) */
    int8_t* wake = (int8_t*)malloc((size_t)(n ? n : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const particle* p = &s->parts[i];
      int8_t w = p->timebin;
      for (int64_t k = 0; k < p->face_count; ++k) {
        const face_t* face = &s->faces[s->conn[p->face_connections_offset + k]];
        int64_t other = (i == face->left) ? face->right : face->left;
        if (other < 0) continue;
        if (!active[other]) continue;
        int8_t cand = (int8_t)(s->parts[other].timebin + TIME_BIN_NEIGHBOUR_MAX_DELTA_BIN);
        if (cand < w) w = cand;
      }
      wake[i] = w;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) timestep_limit(&s->parts[i], wake[i], t);
    free(wake);
  }
  /* reset_extrapolations (hydrodynamics.rs:193-195) */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    if (active[i]) s->parts[i].extrapolations = st_vacuum();
  kick1(s, active);
  free(active);
  *ti_next_out = ti_next;
  return s->error;
}

/* ======================= 1D tessellation ======================= */
/* meshless_voronoi 1D semantics as pinned by test_init_1d (src/space.rs:1453-1491): faces in order of
 * position (left wall, interior faces at generator midpoints, right wall); area 1; the cell volume is
 * the distance between its faces and its centroid their midpoint. Generators must be sorted along x. */
int ora_voronoi_1d(int64_t n, const double* x, double box_x, int periodic, double* area, double* normal,
                   double* centroid, int64_t* left, int64_t* right, double* shift, int64_t* conn, int64_t* offset,
                   int64_t* count, double* volume, double* cell_centroid, double* search_radius,
                   int64_t* n_faces_out) {
  if (n < 1) return ORA_ERR_UNSUPPORTED;
  /* Generators in index order along x; on a periodic line they may be a rotation of that order (a
   * generator that drifted through the box edge keeps its neighbours): exactly one step of the cycle then
   * goes down, and that face carries the periodic shift. */
  int64_t descents = 0;
  for (int64_t i = 0; i + 1 < n; ++i)
    if (!(x[i] < x[i + 1])) ++descents;
  if (!periodic && descents) return ORA_ERR_UNSUPPORTED;
  if (periodic && descents + (x[n - 1] < x[0] ? 0 : 1) != 1) return ORA_ERR_UNSUPPORTED;
  int64_t nf = periodic ? n : n + 1;
  for (int64_t f = 0; f < nf; ++f) {
    area[f] = 1.;
    normal[3 * f + 0] = 1.;
    normal[3 * f + 1] = 0.;
    normal[3 * f + 2] = 0.;
    for (int k = 0; k < 3; ++k) {
      centroid[3 * f + k] = 0.;
      shift[3 * f + k] = 0.;
    }
  }
  /* face k (1 <= k <= n-1) sits between generators k-1 and k (on a rotated periodic line, the face of
   * the downward step joins generator k-1 to the image of generator k one box to the right) */
  for (int64_t k = 1; k < n; ++k) {
    if (x[k - 1] < x[k]) {
      centroid[3 * k] = 0.5 * (x[k - 1] + x[k]);
    } else {
      centroid[3 * k] = 0.5 * (x[k - 1] + (x[k] + box_x));
      shift[3 * k] = box_x;
    }
    left[k] = k - 1;
    right[k] = k;
  }
  if (!periodic) {
    centroid[0] = 0.;
    normal[0] = -1.;
    left[0] = 0;
    right[0] = -1;
    centroid[3 * n] = box_x;
    left[n] = n - 1;
    right[n] = -1;
  } else {
    /* face 0 is the face between generator n-1 (left) and generator 0: the wrap face with the periodic
     * shift, unless a rotation moved the downward step elsewhere */
    left[0] = n - 1;
    right[0] = 0;
    if (x[n - 1] < x[0]) {
      centroid[0] = 0.5 * (x[n - 1] + x[0]);
    } else {
      centroid[0] = 0.5 * (x[n - 1] + (x[0] + box_x));
      shift[0] = box_x;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    double xl, xr;
    if (!periodic) {
      xl = centroid[3 * i];
      xr = centroid[3 * (i + 1)];
    } else {
      /* the cell's left face (face i) is given in cell i-1's frame: one box to the right when shifted */
      xl = centroid[3 * i] - shift[3 * i];
      xr = (i == n - 1) ? centroid[0] : centroid[3 * (i + 1)];
    }
    volume[i] = xr - xl;
    cell_centroid[3 * i] = 0.5 * (xl + xr);
    cell_centroid[3 * i + 1] = 0.;
    cell_centroid[3 * i + 2] = 0.;
    search_radius[i] = 1.5 * volume[i];
    offset[i] = 2 * i;
    count[i] = 2;
    if (!periodic) {
      conn[2 * i] = i;
      conn[2 * i + 1] = i + 1;
    } else {
      conn[2 * i] = i;
      conn[2 * i + 1] = (i + 1) % n;
      if (i == n - 1) { /* ascending face order */
        conn[2 * i] = 0;
        conn[2 * i + 1] = n - 1;
      }
    }
  }
  *n_faces_out = nf;
  return ORA_OK;
}

static void set_mesh(ora_sim* s, const ora_mesh* m, const uint8_t* update_mask) {
  free(s->faces);
  free(s->conn);
  s->nf = m->n_faces;
  s->faces = (face_t*)malloc(sizeof(face_t) * (size_t)(s->nf ? s->nf : 1));
  for (int64_t f = 0; f < s->nf; ++f) {
    face_t* fc = &s->faces[f];
    fc->area = m->area[f];
    fc->normal = v3(m->normal[3 * f], m->normal[3 * f + 1], m->normal[3 * f + 2]);
    fc->centroid = v3(m->centroid[3 * f], m->centroid[3 * f + 1], m->centroid[3 * f + 2]);
    fc->shift = v3(m->shift[3 * f], m->shift[3 * f + 1], m->shift[3 * f + 2]);
    fc->left = m->left[f];
    fc->right = m->right[f];
  }
  s->nconn = m->n_conn;
  s->conn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->nconn ? s->nconn : 1));
  memcpy(s->conn, m->conn, sizeof(int64_t) * (size_t)s->nconn);
  for (int64_t i = 0; i < s->n; ++i) {
    particle* p = &s->parts[i];
    p->face_connections_offset = m->offset[i];
    p->face_count = m->count[i];
    if (!update_mask || update_mask[i]) { /* Particle::update_geometry (part.rs:151-158) */
      p->volume = m->volume[i];
      p->search_radius = m->search_radius[i];
      p->centroid = v3(m->cell_centroid[3 * i], m->cell_centroid[3 * i + 1], m->cell_centroid[3 * i + 2]);
    }
  }
}

/* Space::volume_calculation (space.rs:236-263) for 1D; 2D/3D moving meshes are rebuilt by the host. */
static int mesh_update_1d(ora_sim* s, const uint8_t* mask) {
  if (s->p.dim != 1) return ORA_ERR_UNSUPPORTED;
  int64_t n = s->n;
  int periodic = s->p.boundary == ORA_PERIODIC;
  int64_t nfmax = n + 1;
  double* x = (double*)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) x[i] = s->parts[i].loc.x;
  double* area = (double*)malloc(sizeof(double) * nfmax);
  double* normal = (double*)malloc(sizeof(double) * 3 * nfmax);
  double* centroid = (double*)malloc(sizeof(double) * 3 * nfmax);
  double* shift = (double*)malloc(sizeof(double) * 3 * nfmax);
  int64_t* left = (int64_t*)malloc(sizeof(int64_t) * nfmax);
  int64_t* right = (int64_t*)malloc(sizeof(int64_t) * nfmax);
  int64_t* conn = (int64_t*)malloc(sizeof(int64_t) * 2 * n);
  int64_t* offset = (int64_t*)malloc(sizeof(int64_t) * n);
  int64_t* count = (int64_t*)malloc(sizeof(int64_t) * n);
  double* vol = (double*)malloc(sizeof(double) * n);
  double* cc = (double*)malloc(sizeof(double) * 3 * n);
  double* sr = (double*)malloc(sizeof(double) * n);
  int64_t nf = 0;
  int e = ora_voronoi_1d(n, x, s->p.box[0], periodic, area, normal, centroid, left, right, shift, conn, offset, count,
                         vol, cc, sr, &nf);
  if (!e) {
    ora_mesh m = {nf, area, normal, centroid, left, right, shift, 2 * n, conn, offset, count, vol, cc, sr};
    set_mesh(s, &m, mask);
  }
  free(x); free(area); free(normal); free(centroid); free(shift); free(left); free(right);
  free(conn); free(offset); free(count); free(vol); free(cc); free(sr);
  return e;
}

/* Space::volume_calculation (space.rs:236-263) for 2D/3D through the caller's tessellator. */
static int mesh_update_caller(ora_sim* s, const uint8_t* mask) {
  if (!s->mesher) return ORA_ERR_UNSUPPORTED;
  int64_t n = s->n;
  double* loc = (double*)malloc(sizeof(double) * 3 * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    loc[3 * i] = s->parts[i].loc.x;
    loc[3 * i + 1] = s->parts[i].loc.y;
    loc[3 * i + 2] = s->parts[i].loc.z;
  }
  s->mesh_mask = mask;
  int e = s->mesher(s->mesher_user, s, n, loc, mask);
  s->mesh_mask = NULL;
  free(loc);
  return e ? ORA_ERR_MESH : ORA_OK;
}

void ora_sim_set_mesher(ora_sim* s, ora_mesher mesher, void* user) {
  s->mesher = mesher;
  s->mesher_user = user;
}

void ora_sim_apply_mesh(ora_sim* s, const ora_mesh* mesh) { set_mesh(s, mesh, s->mesh_mask); }

/* ======================= sim / engine ======================= */
ora_sim* ora_sim_new(const ora_params* p, int64_t n, const double* loc, const double* mass, const double* vel,
                     const double* u_int, const ora_mesh* mesh) {
  ora_sim* s = (ora_sim*)calloc(1, sizeof(ora_sim));
  s->p = *p;
  s->eos.gamma = adiabatic_new(p->gamma);
  s->eos.eos = p->eos;
  s->eos.iso_u = p->isothermal_u;
  s->n = n;
  s->parts = (particle*)calloc((size_t)(n ? n : 1), sizeof(particle));
  for (int64_t i = 0; i < n; ++i) { /* Particle::from_ic (part.rs:207-217) */
    particle* q = &s->parts[i];
    vec3 x = v3(loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]);
    vec3 v = v3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    q->loc = x;
    q->conserved = st_make(mass[i], v3_scale(mass[i], v), mass[i] * u_int[i] + 0.5 * mass[i] * v3_len2(v));
  }
  set_mesh(s, mesh, 0);
  /* Particle::first_init (part.rs:200-205) */
  for (int64_t i = 0; i < n; ++i) {
    particle* q = &s->parts[i];
    st5 c = q->conserved;
    if (c.c[0] > 0.) {
      double m_inv = 1. / c.c[0];
      double density = c.c[0] / q->volume;
      vec3 velocity = v3_scale(m_inv, st_vec(c));
      double thermal = c.c[4] - 0.5 * v3_len2(st_vec(c)) * m_inv;
      double pressure = gas_pressure_from_internal_energy(&s->eos, thermal * m_inv, density);
      if (!(density >= 0.) || !(pressure >= 0.)) set_error(s, ORA_ERR_NEG_PRESSURE);
      q->primitives = st_make(density, velocity, pressure);
    } else {
      q->primitives = st_vacuum();
    }
    for (int k = 0; k < 5; ++k)
      if (!isfinite(q->primitives.c[k])) set_error(s, ORA_ERR_NONFINITE);
  }
  /* Engine::new (engine.rs:104-149) */
  s->ts.time_base = p->t_end / (double)MAX_NR_TIMESTEPS;
  s->ts.time_base_inv = 1. / s->ts.time_base;
  s->ts.ti_old = s->ts.ti_current = s->ts.ti_next = 0;
  s->t_current = 0.;
  s->sp_cap = 64;
  s->sp_ti = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)s->sp_cap);
  s->sp_kind = (int*)malloc(sizeof(int) * (size_t)s->sp_cap);
  s->sp_head = s->sp_tail = 0;
  s->sp_ti[s->sp_tail] = 0;
  s->sp_kind[s->sp_tail++] = SP_STEP;
  s->ti_snap = 0;
  s->ti_between_snaps =
      p->dt_snap > 0. ? ora_saturating_f64_to_u64(p->dt_snap * s->ts.time_base_inv) : MAX_NR_TIMESTEPS;
  if (s->ti_between_snaps == 0) s->ti_between_snaps = MAX_NR_TIMESTEPS;
  return s;
}

void ora_sim_free(ora_sim* s) {
  if (!s) return;
  free(s->parts);
  free(s->faces);
  free(s->conn);
  free(s->sp_ti);
  free(s->sp_kind);
  free(s);
}

/* OpenMP threads of the oracle's parallel loops (and of every later OpenMP region in the process):
 * n > 0 threads, n <= 0 the default the process started with.  Returns the count now in effect. */
int ora_sim_set_threads(int n) {
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  extern int omp_get_max_threads(void);
  static int initial = 0;
  if (initial == 0) initial = omp_get_max_threads();
  omp_set_num_threads(n > 0 ? n : initial);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

static void sp_push(ora_sim* s, uint64_t ti, int kind) {
  if (s->sp_tail == s->sp_cap) {
    int64_t live = s->sp_tail - s->sp_head;
    memmove(s->sp_ti, s->sp_ti + s->sp_head, sizeof(uint64_t) * (size_t)live);
    memmove(s->sp_kind, s->sp_kind + s->sp_head, sizeof(int) * (size_t)live);
    s->sp_head = 0;
    s->sp_tail = live;
    if (s->sp_tail == s->sp_cap) {
      s->sp_cap *= 2;
      s->sp_ti = (uint64_t*)realloc(s->sp_ti, sizeof(uint64_t) * (size_t)s->sp_cap);
      s->sp_kind = (int*)realloc(s->sp_kind, sizeof(int) * (size_t)s->sp_cap);
    }
  }
  s->sp_ti[s->sp_tail] = ti;
  s->sp_kind[s->sp_tail++] = kind;
}

/* utils.rs:15-43 */
static int box_contains(const double* box, vec3 pos, int dim) {
  int c = 1;
  for (int i = 0; i < dim; ++i) c &= (v3_get(pos, i) >= 0. && v3_get(pos, i) < box[i]);
  return c;
}

/* Space::drift (space.rs:833-854) with Particle::drift (part.rs:107-119), box_wrap / box_reflect
 * (part.rs:449-469). */
static int space_drift(ora_sim* s, double dt) {
  int dim = s->p.dim;
  int err = 0;
  for (int64_t i = 0; i < s->n; ++i) {
    particle* p = &s->parts[i];
    for (int k = 0; k < dim; ++k) {
      v3_set(&p->loc, k, v3_get(p->loc, k) + v3_get(p->v, k) * dt);
      v3_set(&p->centroid, k, v3_get(p->centroid, k) + v3_get(p->v, k) * dt);
      v3_set(&p->gradients_centroid, k, v3_get(p->gradients_centroid, k) + v3_get(p->v, k) * dt);
    }
    if (!v3_finite(p->loc)) err = ORA_ERR_DRIFT;
  }
  for (int64_t i = 0; i < s->n; ++i) {
    particle* p = &s->parts[i];
    if (box_contains(s->p.box, p->loc, dim)) continue;
    if (s->p.boundary == ORA_REFLECTIVE) {
      vec3 old = p->loc;
      for (int k = 0; k < dim; ++k) {
        double x = v3_get(p->loc, k);
        if (x < 0.) x -= 2. * x;
        if (x > s->p.box[k]) x -= 2. * (x - s->p.box[k]);
        v3_set(&p->loc, k, x);
      }
      if (!box_contains(s->p.box, p->loc, dim)) err = ORA_ERR_DRIFT;
      vec3 sh = v3_sub(p->loc, old);
      if (v3_len2(sh) > 0.) {
        vec3 nrm = v3_scale(1. / v3_len(sh), sh);
        p->primitives = prim_reflect(p->primitives, nrm);
        reflect_gradients(p, nrm);
      }
    } else if (s->p.boundary == ORA_PERIODIC) {
      vec3 old = p->loc;
      for (int k = 0; k < dim; ++k) {
        double x = v3_get(p->loc, k);
        while (x < 0.) x += s->p.box[k];
        while (x >= s->p.box[k]) x -= s->p.box[k];
        v3_set(&p->loc, k, x);
      }
      vec3 sh = v3_sub(p->loc, old);
      p->gradients_centroid = v3_add(p->gradients_centroid, sh);
      p->centroid = v3_add(p->centroid, sh);
    } else {
      err = ORA_ERR_UNSUPPORTED; /* particle removal for open/vacuum boundaries is out of scope */
    }
  }
  return err;
}

/* Engine::run / step / queue_step / update_ti_next (engine.rs:152-268) */
int64_t ora_sim_advance(ora_sim* s, int64_t max_steps) {
  int64_t done = 0;
  while (s->t_current < s->p.t_end && done < max_steps && !s->error) {
    if (s->sp_head == s->sp_tail) {
      s->error = ORA_ERR_UNSUPPORTED;
      break;
    }
    uint64_t ti = s->sp_ti[s->sp_head];
    int kind = s->sp_kind[s->sp_head];
    s->sp_head++;
    if (ti < s->ts.ti_current) {
      s->error = ORA_ERR_UNSUPPORTED;
      break;
    }
    uint64_t dti = ti - s->ts.ti_current;
    if (dti > 0) {
      double dt = dt_from_dti(&s->ts, dti);
      int e = space_drift(s, dt);
      if (e) set_error(s, e);
      s->ts.ti_current += dti;
      s->t_current = make_timestep(s->ts.ti_current, s->ts.time_base);
    }
    if (kind == SP_DUMP) continue; /* snapshots are written by the host harness, not the oracle */
    uint64_t ti_next = 0;
    ora_sim_runner_step(s, &ti_next);
    if (s->error) break;
    /* queue_step */
    while (s->ti_snap < ti_next) {
      sp_push(s, s->ti_snap, SP_DUMP);
      s->ti_snap += s->ti_between_snaps;
    }
    sp_push(s, ti_next, SP_STEP);
    /* update_ti_next */
    s->ts.ti_old = s->ts.ti_next;
    s->ts.ti_next = ti_next;
    s->step_count++;
    done++;
  }
  if (s->error) return s->error;
  return done;
}

void ora_sim_engine_info(const ora_sim* s, ora_engine_info* out) {
  out->ti_old = s->ts.ti_old;
  out->ti_current = s->ts.ti_current;
  out->ti_next = s->ts.ti_next;
  out->time_base = s->ts.time_base;
  out->time_base_inv = s->ts.time_base_inv;
  out->t_current = s->t_current;
  out->t_end = s->p.t_end;
  out->step_count = s->step_count;
  out->error = s->error;
  out->last_n_active = s->last_n_active;
  out->last_n_face_active = s->last_n_face_active;
}

void ora_sim_set_engine_info(ora_sim* s, uint64_t ti_old, uint64_t ti_current, uint64_t ti_next) {
  s->ts.ti_old = ti_old;
  s->ts.ti_current = ti_current;
  s->ts.ti_next = ti_next;
  s->t_current = make_timestep(ti_current, s->ts.time_base);
}

void ora_sim_counters(const ora_sim* s, ora_counters* out) { *out = s->cnt; }
int64_t ora_sim_n_cells(const ora_sim* s) { return s->n; }
int64_t ora_sim_n_faces(const ora_sim* s) { return s->nf; }

void ora_sim_get_mesh(const ora_sim* s, double* area, double* normal, double* centroid, int64_t* left,
                      int64_t* right, double* shift, int64_t* conn, int64_t* offset, int64_t* count) {
  for (int64_t f = 0; f < s->nf; ++f) {
    const face_t* fc = &s->faces[f];
    if (area) area[f] = fc->area;
    if (normal) { normal[3 * f] = fc->normal.x; normal[3 * f + 1] = fc->normal.y; normal[3 * f + 2] = fc->normal.z; }
    if (centroid) {
      centroid[3 * f] = fc->centroid.x; centroid[3 * f + 1] = fc->centroid.y; centroid[3 * f + 2] = fc->centroid.z;
    }
    if (left) left[f] = fc->left;
    if (right) right[f] = fc->right;
    if (shift) { shift[3 * f] = fc->shift.x; shift[3 * f + 1] = fc->shift.y; shift[3 * f + 2] = fc->shift.z; }
  }
  if (conn) memcpy(conn, s->conn, sizeof(int64_t) * (size_t)s->nconn);
  for (int64_t i = 0; i < s->n; ++i) {
    if (offset) offset[i] = s->parts[i].face_connections_offset;
    if (count) count[i] = s->parts[i].face_count;
  }
}

#define GET3(arr, v)               \
  do {                             \
    (arr)[i] = (v).x;              \
    (arr)[n + i] = (v).y;          \
    (arr)[2 * n + i] = (v).z;      \
  } while (0)
#define SET3(v, arr) (v) = v3((arr)[i], (arr)[n + i], (arr)[2 * n + i])

void ora_sim_get_fields(const ora_sim* s, ora_fields* f) {
  int64_t n = s->n;
  for (int64_t i = 0; i < n; ++i) {
    const particle* p = &s->parts[i];
    for (int k = 0; k < 5; ++k) {
      if (f->prim) f->prim[k * n + i] = p->primitives.c[k];
      if (f->extrap) f->extrap[k * n + i] = p->extrapolations.c[k];
      if (f->cons) f->cons[k * n + i] = p->conserved.c[k];
      if (f->fluxes) f->fluxes[k * n + i] = p->fluxes.c[k];
      if (f->grad) {
        f->grad[(k * 3 + 0) * n + i] = p->gradients.g[k].x;
        f->grad[(k * 3 + 1) * n + i] = p->gradients.g[k].y;
        f->grad[(k * 3 + 2) * n + i] = p->gradients.g[k].z;
      }
    }
    if (f->gc) GET3(f->gc, p->gradients_centroid);
    if (f->loc) GET3(f->loc, p->loc);
    if (f->centroid) GET3(f->centroid, p->centroid);
    if (f->v) GET3(f->v, p->v);
    if (f->v_rel) GET3(f->v_rel, p->v_rel);
    if (f->volume) f->volume[i] = p->volume;
    if (f->maor) f->maor[i] = p->max_a_over_r;
    if (f->msv) f->msv[i] = p->max_signal_velocity;
    if (f->dt) f->dt[i] = p->dt;
    if (f->timebin) f->timebin[i] = p->timebin;
  }
}

void ora_sim_set_fields(ora_sim* s, const ora_fields* f) {
  int64_t n = s->n;
  for (int64_t i = 0; i < n; ++i) {
    particle* p = &s->parts[i];
    for (int k = 0; k < 5; ++k) {
      if (f->prim) p->primitives.c[k] = f->prim[k * n + i];
      if (f->extrap) p->extrapolations.c[k] = f->extrap[k * n + i];
      if (f->cons) p->conserved.c[k] = f->cons[k * n + i];
      if (f->fluxes) p->fluxes.c[k] = f->fluxes[k * n + i];
      if (f->grad)
        p->gradients.g[k] = v3(f->grad[(k * 3 + 0) * n + i], f->grad[(k * 3 + 1) * n + i], f->grad[(k * 3 + 2) * n + i]);
    }
    if (f->gc) SET3(p->gradients_centroid, f->gc);
    if (f->loc) SET3(p->loc, f->loc);
    if (f->centroid) SET3(p->centroid, f->centroid);
    if (f->v) SET3(p->v, f->v);
    if (f->v_rel) SET3(p->v_rel, f->v_rel);
    if (f->volume) p->volume = f->volume[i];
    if (f->maor) p->max_a_over_r = f->maor[i];
    if (f->msv) p->max_signal_velocity = f->msv[i];
    if (f->dt) p->dt = f->dt[i];
    if (f->timebin) p->timebin = f->timebin[i];
  }
}

/* space.rs:1248-1280 (mass and energy including pending fluxes), extended with momentum. */
void ora_sim_conservation(const ora_sim* s, double* out5) {
  for (int k = 0; k < 5; ++k) out5[k] = 0.;
  for (int64_t i = 0; i < s->n; ++i)
    for (int k = 0; k < 5; ++k) out5[k] += s->parts[i].conserved.c[k] + s->parts[i].fluxes.c[k];
}

/* ======================= single functions, for the tests that pin them ======================= */
static gaslaw ideal_gas(double gamma) {
  gaslaw g;
  g.gamma = adiabatic_new(gamma);
  g.eos = ORA_IDEAL;
  g.iso_u = 0.;
  return g;
}

int ora_t_c2p(double gamma, const double* cons5, double volume, double* prim5) {
  gaslaw g = ideal_gas(gamma);
  particle p;
  memset(&p, 0, sizeof(p));
  for (int k = 0; k < 5; ++k) p.conserved.c[k] = cons5[k];
  p.volume = volume;
  int e = convert_conserved_to_primitive(&p, &g);
  for (int k = 0; k < 5; ++k) prim5[k] = p.primitives.c[k];
  return e;
}

void ora_t_time_extrapolations(double gamma, const double* prim5, const double* grad15, const double* v_rel3,
                               double dt, double* out5) {
  gaslaw g = ideal_gas(gamma);
  particle p;
  memset(&p, 0, sizeof(p));
  for (int k = 0; k < 5; ++k) {
    p.primitives.c[k] = prim5[k];
    p.gradients.g[k] = v3(grad15[3 * k], grad15[3 * k + 1], grad15[3 * k + 2]);
  }
  p.v_rel = v3(v_rel3[0], v_rel3[1], v_rel3[2]);
  st5 r = time_extrapolations(&p, dt, &g);
  for (int k = 0; k < 5; ++k) out5[k] = r.c[k];
}

double ora_t_radius(double volume, int dim) {
  particle p;
  memset(&p, 0, sizeof(p));
  p.volume = volume;
  return radius(&p, dim);
}

/* in: cons[5], prim[5], grad[15] (component-major xyz per component), loc[3], gc[3], centroid[3], volume,
 * msv, maor; out: dt, v[3], v_rel[3], msv after the call. Returns the error flag of set_particle_velocity. */
int ora_t_particle_timestep(double gamma, double cfl, int motion, int dim, const double* in, double* out) {
  gaslaw g = ideal_gas(gamma);
  particle p;
  memset(&p, 0, sizeof(p));
  for (int k = 0; k < 5; ++k) {
    p.conserved.c[k] = in[k];
    p.primitives.c[k] = in[5 + k];
    p.gradients.g[k] = v3(in[10 + 3 * k], in[11 + 3 * k], in[12 + 3 * k]);
  }
  p.loc = v3(in[25], in[26], in[27]);
  p.gradients_centroid = v3(in[28], in[29], in[30]);
  p.centroid = v3(in[31], in[32], in[33]);
  p.volume = in[34];
  p.max_signal_velocity = in[35];
  p.max_a_over_r = in[36];
  int err = 0;
  out[0] = particle_timestep(&p, cfl, motion, &g, dim, &err);
  out[1] = p.v.x;
  out[2] = p.v.y;
  out[3] = p.v.z;
  out[4] = p.v_rel.x;
  out[5] = p.v_rel.y;
  out[6] = p.v_rel.z;
  out[7] = p.max_signal_velocity;
  return err;
}

int8_t ora_t_timestep_limit(double dt, int8_t timebin, int8_t wakeup, uint64_t ti_current, double time_base,
                            double* dt_out) {
  particle p;
  memset(&p, 0, sizeof(p));
  p.dt = dt;
  p.timebin = timebin;
  tsinfo t;
  memset(&t, 0, sizeof(t));
  t.ti_current = ti_current;
  t.time_base = time_base;
  t.time_base_inv = 1. / time_base;
  timestep_limit(&p, wakeup, &t);
  *dt_out = p.dt;
  return p.timebin;
}

void ora_t_face_velocity(const double* loc_l, const double* loc_r, const double* shift, const double* v_l,
                         const double* v_r, const double* fc, double* out3) {
  vec3 w = face_velocity(v3(loc_l[0], loc_l[1], loc_l[2]), v3(loc_r[0], loc_r[1], loc_r[2]),
                         v3(shift[0], shift[1], shift[2]), v3(v_l[0], v_l[1], v_l[2]), v3(v_r[0], v_r[1], v_r[2]),
                         v3(fc[0], fc[1], fc[2]));
  out3[0] = w.x;
  out3[1] = w.y;
  out3[2] = w.z;
}

/* gradient_estimate + gradient_limit + gradient_apply (hydrodynamics.rs:67-143) on the sim's current
 * primitives, for the cells flagged in `active`. */
void ora_phase_gradients(ora_sim* s, const uint8_t* active) { gradients_phase(s, active); }
