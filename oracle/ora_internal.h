/* Internal types of the CPU oracle (test infrastructure only; see oracle.h). */
#ifndef HYDRO_ORA_INTERNAL_H
#define HYDRO_ORA_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include "oracle.h"

/* glam::DVec3 (glam 0.27 scalar implementation). */
typedef struct { double x, y, z; } vec3;

static inline vec3 v3(double x, double y, double z) { vec3 r = {x, y, z}; return r; }
static inline vec3 v3_add(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline vec3 v3_sub(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
/* f64 * DVec3 and DVec3 * f64 (componentwise) */
static inline vec3 v3_scale(double s, vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline vec3 v3_neg(vec3 a) { return v3(-a.x, -a.y, -a.z); }
/* DVec3::dot: (x*x' + y*y') + z*z' */
static inline double v3_dot(vec3 a, vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline double v3_len2(vec3 a) { return v3_dot(a, a); }
static inline double v3_len(vec3 a) { return sqrt(v3_dot(a, a)); }
/* DVec3::cross */
static inline vec3 v3_cross(vec3 a, vec3 b) {
  return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline int v3_finite(vec3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static inline double v3_get(vec3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void v3_set(vec3* a, int i, double v) { if (i == 0) a->x = v; else if (i == 1) a->y = v; else a->z = v; }
static const vec3 V3_ZERO = {0., 0., 0.};

/* State<T>(f64, DVec3, f64) (src/physical_quantities.rs:16). c[0]=rho|mass, c[1..3]=v|p, c[4]=P|E */
typedef struct { double c[5]; } st5;
static inline st5 st_make(double a, vec3 v, double b) { st5 s = {{a, v.x, v.y, v.z, b}}; return s; }
static inline vec3 st_vec(st5 s) { return v3(s.c[1], s.c[2], s.c[3]); }
static inline st5 st_add(st5 a, st5 b) { st5 r; for (int i = 0; i < 5; ++i) r.c[i] = a.c[i] + b.c[i]; return r; }
static inline st5 st_sub(st5 a, st5 b) { st5 r; for (int i = 0; i < 5; ++i) r.c[i] = a.c[i] - b.c[i]; return r; }
static inline st5 st_scale(double s, st5 a) { st5 r; for (int i = 0; i < 5; ++i) r.c[i] = s * a.c[i]; return r; }
static inline st5 st_vacuum(void) { st5 r = {{0., 0., 0., 0., 0.}}; return r; }

/* Gradients<T>(DVec3, [DVec3;3], DVec3) (src/physical_quantities.rs:115) */
typedef struct { vec3 g[5]; } grad5;

/* AdiabaticIndex (src/gas_law.rs:1-54): the derived constants are recomputed on each call in
 * the reference; computing them once gives bit-identical values. */
typedef struct {
  double gamma, gamma_inv, odgm1, odgp1;
  double gp1dg, gm1d2g, gm1dgp1, tdgm1, tdgp1;
} adiabatic;
static inline adiabatic adiabatic_new(double g) {
  adiabatic a;
  a.gamma = g;
  a.gamma_inv = 1. / g;
  a.odgm1 = 1. / (g - 1.);
  a.odgp1 = 1. / (g + 1.);
  a.gp1dg = (g + 1.) * a.gamma_inv;
  a.gm1d2g = 0.5 * (g - 1.) * a.gamma_inv;
  a.gm1dgp1 = (g - 1.) * a.odgp1;
  a.tdgm1 = 2. * a.odgm1;
  a.tdgp1 = 2. * a.odgp1;
  return a;
}

typedef struct {
  adiabatic gamma;
  int eos;
  double iso_u;
} gaslaw;

/* GasLaw (src/gas_law.rs:81-106) */
static inline double gas_internal_energy_from_pressure(const gaslaw* g, double p, double rho_inv) {
  if (g->eos == ORA_IDEAL) return p * rho_inv * g->gamma.odgm1;
  return g->iso_u;
}
static inline double gas_pressure_from_internal_energy(const gaslaw* g, double u, double rho) {
  if (g->eos == ORA_IDEAL) return (g->gamma.gamma - 1.) * u * rho;
  return (g->gamma.gamma - 1.) * g->iso_u * rho;
}
static inline double gas_sound_speed(const gaslaw* g, double p, double rho_inv) {
  if (g->eos == ORA_IDEAL) return sqrt(g->gamma.gamma * p * rho_inv);
  return sqrt(g->iso_u * g->gamma.gamma * (g->gamma.gamma - 1.));
}

/* Rust f64::min / f64::max ignore a NaN operand, like C fmin/fmax. */
static inline double rmin(double a, double b) { return fmin(a, b); }
static inline double rmax(double a, double b) { return fmax(a, b); }

/* Riemann */
typedef struct { double rho_l, rho_r, u, p; } star_values;

int ora_riemann_solve(const ora_riemann* rs, st5 left, st5 right, vec3 iv, vec3 n, const gaslaw* eos,
                      st5* flux, ora_counters* cnt);
st5 ora_pairwise(st5 pl, st5 pr, st5 pdash, double dx_fac, ora_counters* cnt);
void ora_check_physical(st5* s, ora_counters* cnt);

#endif
