/* CPU oracle (test infrastructure only): Riemann solvers and pairwise limiter.
 * Restates src/riemann_solver.rs, src/riemann_solver/{exact,hllc,pvrs,airs,tsrs,trrs,vacuum,
 * linear_advection}.rs, src/gradients.rs:105-152, src/physical_quantities.rs:228-267. */
#include "ora_internal.h"

/* State<Primitive>::check_physical (src/physical_quantities.rs:245-267) */
void ora_check_physical(st5* s, ora_counters* cnt) {
  if (s->c[0] < 0.) {
    if (cnt) cnt->check_physical_resets++;
    *s = st_vacuum();
  }
  if (s->c[4] < 0.) {
    if (cnt) cnt->check_physical_resets++;
    *s = st_vacuum();
  }
}

/* flux_from_half_state (src/riemann_solver.rs:46-76) */
static st5 flux_from_half_state(st5 half, vec3 iv, vec3 n, const adiabatic* g) {
  vec3 hv = st_vec(half);
  double rho = half.c[0], P = half.c[4];
  vec3 v_tot = v3_add(iv, hv);
  vec3 f0 = v3_scale(rho, hv);
  vec3 f1 = v3_scale(rho * v_tot.x, hv);
  f1.x += P;
  vec3 f2 = v3_scale(rho * v_tot.y, hv);
  f2.y += P;
  vec3 f3 = v3_scale(rho * v_tot.z, hv);
  f3.z += P;
  double roe = P * g->odgm1 + 0.5 * rho * v3_len2(v_tot);
  vec3 f4 = v3_add(v3_scale(roe, hv), v3_scale(P, v_tot));
  return st_make(v3_dot(f0, n), v3(v3_dot(f1, n), v3_dot(f2, n), v3_dot(f3, n)), v3_dot(f4, n));
}

/* ---- generic star-state sampling (src/riemann_solver.rs:96-256) ---- */
static double shock_speed(double v, double a, double pdps, const adiabatic* g) {
  return v - a * sqrt(0.5 * g->gp1dg * pdps + g->gm1d2g);
}
static double rarefaction_head_speed(double v, double a) { return v - a; }
static double rarefaction_tail_speed(double u, double a, double pdps, const adiabatic* g) {
  return u - a * pow(pdps, g->gm1d2g);
}
static st5 sample_rarefaction_fan(st5 s, double a, double v, vec3 n, const adiabatic* g) {
  double v_half = g->tdgp1 * (a + 0.5 * (g->gamma - 1.) * v);
  double base = g->tdgp1 + g->gm1dgp1 / a * v;
  return st_make(s.c[0] * pow(base, g->tdgm1), v3_add(st_vec(s), v3_scale(v_half - v, n)),
                 s.c[4] * pow(base, g->gamma * g->tdgm1));
}
static st5 sample_middle_state(double rho, double u, double p, st5 s, double v, vec3 n) {
  return st_make(rho, v3_add(st_vec(s), v3_scale(u - v, n)), p);
}
static st5 sample_left_rarefaction_wave(double rho, double u, double p, st5 left, double v, double a, vec3 n,
                                        const adiabatic* g) {
  if (rarefaction_head_speed(v, a) < 0.) {
    if (rarefaction_tail_speed(u, a, p / left.c[4], g) > 0.) return sample_rarefaction_fan(left, a, v, n, g);
    return sample_middle_state(rho, u, p, left, v, n);
  }
  return left;
}
static st5 sample_right_rarefaction_wave(double rho, double u, double p, st5 right, double v, double a, vec3 n,
                                         const adiabatic* g) {
  if (rarefaction_head_speed(v, -a) > 0.) {
    if (rarefaction_tail_speed(u, -a, p / right.c[4], g) < 0.) return sample_rarefaction_fan(right, -a, v, n, g);
    return sample_middle_state(rho, u, p, right, v, n);
  }
  return right;
}
static st5 sample_left_shock_wave(double rho, double u, double p, st5 left, double v, double a, vec3 n,
                                  const adiabatic* g) {
  if (shock_speed(v, a, p / left.c[4], g) < 0.) return sample_middle_state(rho, u, p, left, v, n);
  return left;
}
static st5 sample_right_shock_wave(double rho, double u, double p, st5 right, double v, double a, vec3 n,
                                   const adiabatic* g) {
  if (shock_speed(v, -a, p / right.c[4], g) > 0.) return sample_middle_state(rho, u, p, right, v, n);
  return right;
}
static st5 sample_from_star(star_values star, st5 left, st5 right, double v_l, double v_r, double a_l, double a_r,
                            vec3 n, const adiabatic* g) {
  if (star.u < 0.) {
    if (star.p > right.c[4]) return sample_right_shock_wave(star.rho_r, star.u, star.p, right, v_r, a_r, n, g);
    return sample_right_rarefaction_wave(star.rho_r, star.u, star.p, right, v_r, a_r, n, g);
  } else if (star.p > left.c[4]) {
    return sample_left_shock_wave(star.rho_l, star.u, star.p, left, v_l, a_l, n, g);
  }
  return sample_left_rarefaction_wave(star.rho_l, star.u, star.p, left, v_l, a_l, n, g);
}

/* ---- vacuum (src/riemann_solver/vacuum.rs) ---- */
static int is_vacuum(st5 left, st5 right, double a_l, double a_r, double v_r_m_v_l, const adiabatic* g) {
  if (left.c[0] == 0. || right.c[0] == 0.) return 1;
  if (g->tdgm1 * (a_l + a_r) <= v_r_m_v_l) return 1;
  return 0;
}
static st5 sample_right_vacuum(st5 left, double v_l, double a_l, double s_l, vec3 n, const adiabatic* g) {
  if (0. <= v_l - a_l) return left;
  if (0. < s_l) return sample_rarefaction_fan(left, a_l, v_l, n, g);
  return st_vacuum();
}
static st5 sample_left_vacuum(st5 right, double v_r, double a_r, double s_r, vec3 n, const adiabatic* g) {
  if (0. <= s_r) return st_vacuum();
  if (0. < v_r + a_r) return sample_rarefaction_fan(right, -a_r, v_r, n, g);
  return right;
}
static st5 sample_vacuum_creation(st5 left, st5 right, double v_l, double v_r, double a_l, double a_r, vec3 n,
                                  const adiabatic* g) {
  double s_l = v_l + g->tdgm1 * a_l;
  double s_r = v_r - g->tdgm1 * a_r;
  if (0. <= s_l) return sample_right_vacuum(left, v_l, a_l, s_l, n, g);
  if (0. < s_r) return st_vacuum();
  return sample_left_vacuum(right, v_r, a_r, s_r, n, g);
}
/* VacuumRiemannSolver::sample (vacuum.rs:116-138) */
static st5 vacuum_sample(st5 left, st5 right, double v_l, double v_r, double a_l, double a_r, vec3 n,
                         const adiabatic* g) {
  if (right.c[0] == 0.) {
    double s_l = v_l + g->tdgm1 * a_l;
    return sample_right_vacuum(left, v_l, a_l, s_l, n, g);
  } else if (left.c[0] == 0.) {
    double s_r = v_r - g->tdgm1 * a_r;
    return sample_left_vacuum(right, v_r, a_r, s_r, n, g);
  }
  return sample_vacuum_creation(left, right, v_l, v_r, a_l, a_r, n, g);
}

/* ---- exact (src/riemann_solver/exact.rs) ---- */
static double ex_fb(double p, st5 s, double a, const adiabatic* g) {
  if (p > s.c[4]) {
    double cap_a = g->tdgp1 / s.c[0];
    double cap_b = g->gm1dgp1 * s.c[4];
    return (p - s.c[4]) * sqrt(cap_a / (p + cap_b));
  }
  return g->tdgm1 * a * (pow(p / s.c[4], g->gm1d2g) - 1.);
}
static double ex_f(double p, st5 l, st5 r, double v_l, double v_r, double a_l, double a_r, const adiabatic* g) {
  return ex_fb(p, l, a_l, g) + ex_fb(p, r, a_r, g) + (v_r - v_l);
}
/* exact.rs:37-45: the rarefaction branch uses the exponent -(gamma-1)/(2 gamma), as in the reference. */
static double ex_fprimeb(double p, st5 s, double a, const adiabatic* g) {
  if (p > s.c[4]) {
    double cap_a = g->tdgp1 / s.c[0];
    double cap_b = g->gm1dgp1 * s.c[4];
    return (1. - 0.5 * (p - s.c[4]) / (cap_b + p)) * sqrt(cap_a / (p + cap_b));
  }
  return 1. / s.c[0] / a * pow(p / s.c[4], -g->gm1d2g);
}
static double ex_fprime(double p, st5 l, st5 r, double a_l, double a_r, const adiabatic* g) {
  return ex_fprimeb(p, l, a_l, g) + ex_fprimeb(p, r, a_r, g);
}
static double ex_gb(double p, st5 s, const adiabatic* g) {
  double cap_a = g->tdgp1 / s.c[0];
  double cap_b = g->gm1dgp1 * s.c[4];
  return sqrt(cap_a / (p + cap_b));
}
static double ex_guess_p(st5 l, st5 r, double v_l, double v_r, double a_l, double a_r, const adiabatic* g) {
  double p_min = rmin(l.c[4], r.c[4]);
  double p_max = rmax(l.c[4], r.c[4]);
  double q_max = p_max / p_min;
  double ppv = 0.5 * (l.c[4] + r.c[4]) - 0.125 * (v_r - v_l) * (l.c[0] + r.c[0]) * (a_l + a_r);
  ppv = rmax(ppv, 1e-8);
  double p_guess;
  if (q_max <= 2. && p_min <= ppv && ppv <= p_max) {
    p_guess = ppv;
  } else if (ppv < p_min) {
    double base = (a_l + a_r - 0.5 * (g->gamma - 1.) * (v_r - v_l)) /
                  (a_l / pow(l.c[4], g->gm1d2g) + a_r / pow(r.c[4], g->gm1d2g));
    p_guess = pow(base, g->gamma * g->tdgm1);
  } else {
    p_guess = (ex_gb(ppv, l, g) * l.c[4] + ex_gb(ppv, r, g) * r.c[4] - v_r + v_l) /
              (ex_gb(ppv, l, g) + ex_gb(ppv, r, g));
  }
  return rmax(p_guess, 1e-8);
}
static int ex_solve_brent(double lower, double upper, double low_f, double up_f, double tol, st5 l, st5 r,
                          double v_l, double v_r, double a_l, double a_r, const adiabatic* g, double* out) {
  double a = lower, b = upper, c, d = INFINITY;
  double fa = low_f, fb = up_f, fc;
  double s, fs;
  if (fa * fb > 0.) return ORA_ERR_BRENT;
  if (fabs(fa) < fabs(fb)) {
    double t = a; a = b; b = t;
    t = fa; fa = fb; fb = t;
  }
  c = a;
  fc = fa;
  int mflag = 1;
  while (fb != 0. && fabs(a - b) > tol * 0.5 * (a + b)) {
    if (fa != fc && fb != fc) {
      s = a * fb * fc / (fa - fb) / (fa - fc) + b * fa * fc / (fb - fa) / (fb - fc) +
          c * fa * fb / (fc - fa) / (fc - fb);
    } else {
      s = b - fb * (b - a) / (fb - fa);
    }
    double tmp = 0.25 * (3. * a + b);
    if (!((s > tmp && s < b) || (s < tmp && s > b)) || (mflag && fabs(s - b) >= (0.5 * fabs(b - c))) ||
        (!mflag && fabs(s - b) >= (0.5 * fabs(c - d))) || (mflag && fabs(b - c) < 0.5 * tol * (b + c)) ||
        (!mflag && fabs(c - d) < 0.5 * tol * (c + d))) {
      s = 0.5 * (a + b);
      mflag = 1;
    } else {
      mflag = 0;
    }
    fs = ex_f(s, l, r, v_l, v_r, a_l, a_r, g);
    d = c;
    c = b;
    fc = fb;
    if (fa * fs < 0.) {
      b = s;
      fb = fs;
    } else {
      a = s;
      fa = fs;
    }
    if (fabs(fa) < fabs(fb)) {
      double t = a; a = b; b = t;
      t = fa; fa = fb; fb = t;
    }
  }
  *out = b;
  return ORA_OK;
}
static double shock_middle_density(double pdps, st5 s, const adiabatic* g) {
  return s.c[0] * (pdps + g->gm1dgp1) / (g->gm1dgp1 * pdps + 1.);
}
static double rarefaction_middle_density(double pdps, st5 s, const adiabatic* g) {
  return s.c[0] * pow(pdps, 1. / g->gamma);
}
static double middle_density(double p, st5 s, const adiabatic* g) {
  double pdps = p / s.c[4];
  if (pdps > 1.) return shock_middle_density(pdps, s, g);
  return rarefaction_middle_density(pdps, s, g);
}
/* exact.rs:214-238 */
static star_values star_state_from_pstar(double pstar, st5 l, st5 r, double v_l, double v_r, double a_l, double a_r,
                                         const adiabatic* g) {
  star_values sv;
  sv.u = 0.5 * (v_l + v_r) + 0.5 * (ex_fb(pstar, r, a_r, g) - ex_fb(pstar, l, a_l, g));
  sv.rho_l = middle_density(pstar, l, g);
  sv.rho_r = middle_density(pstar, r, g);
  sv.p = pstar;
  return sv;
}
/* exact.rs:242-290 */
static int exact_star(st5 l, st5 r, double v_l, double v_r, double a_l, double a_r, const adiabatic* g,
                      star_values* out, ora_counters* cnt) {
  double p = 0.;
  double p_guess = ex_guess_p(l, r, v_l, v_r, a_l, a_r, g);
  double fp = ex_f(p, l, r, v_l, v_r, a_l, a_r, g);
  double fp_guess = ex_f(p_guess, l, r, v_l, v_r, a_l, a_r, g);
  if (fp * fp_guess >= 0.) {
    int counter = 0;
    while (fabs(p - p_guess) > 1e-6 * 0.5 * (p + p_guess) && fp_guess < 0.0) {
      p = p_guess;
      p_guess = p_guess - fp_guess / ex_fprime(p_guess, l, r, a_l, a_r, g);
      fp_guess = ex_f(p_guess, l, r, v_l, v_r, a_l, a_r, g);
      counter += 1;
      if (counter > 1000) return ORA_ERR_NEWTON;
    }
  }
  if (fabs(p - p_guess) > 1e-6 * 0.5 * (p + p_guess) && fp_guess > 0.) {
    p = 0.;
    fp = ex_f(p, l, r, v_l, v_r, a_l, a_r, g);
    if (cnt) cnt->brent_calls++;
    int e = ex_solve_brent(p, p_guess, fp, fp_guess, 1e-6, l, r, v_l, v_r, a_l, a_r, g, &p);
    if (e) return e;
  } else {
    p = p_guess;
  }
  *out = star_state_from_pstar(p, l, r, v_l, v_r, a_l, a_r, g);
  return ORA_OK;
}

/* ---- PVRS (src/riemann_solver/pvrs.rs) ---- */
static double pv_p_star(double rho_bar, double p_bar, double a_bar, double v_l, double v_r) {
  return p_bar + 0.5 * (v_l - v_r) * rho_bar * a_bar;
}
static star_values pvrs_star(st5 l, st5 r, double v_l, double v_r, double a_l, double a_r) {
  double rho_bar = 0.5 * (l.c[0] + r.c[0]);
  double p_bar = 0.5 * (l.c[4] + r.c[4]);
  double a_bar = 0.5 * (a_l + a_r);
  star_values sv;
  sv.p = pv_p_star(rho_bar, p_bar, a_bar, v_l, v_r);
  sv.u = 0.5 * ((v_l + v_r) + (l.c[4] - r.c[4]) / (rho_bar * a_bar));
  sv.rho_l = l.c[0] + (v_l - sv.u) * rho_bar / a_bar;
  sv.rho_r = r.c[0] + (sv.u - v_r) * rho_bar / a_bar;
  return sv;
}

/* ---- TSRS / TRRS (src/riemann_solver/{tsrs,trrs}.rs) ---- */
static double ts_g(double p, double rho_s, double p_s, const adiabatic* g) {
  double a = g->tdgp1 / rho_s;
  double b = g->gm1dgp1 * p_s;
  return sqrt(a / (p + b));
}
static star_values tsrs_star(st5 l, st5 r, double v_l, double v_r, double a_l, double a_r, const adiabatic* g) {
  double p_guess = pv_p_star(0.5 * (l.c[0] + r.c[0]), 0.5 * (l.c[4] + r.c[4]), 0.5 * (a_l + a_r), v_l, v_r);
  double g_l = ts_g(p_guess, l.c[0], l.c[4], g);
  double g_r = ts_g(p_guess, r.c[0], r.c[4], g);
  double pstar = (g_l * l.c[4] + g_r * r.c[4] - (v_r - v_l)) / (g_l + g_r);
  return star_state_from_pstar(pstar, l, r, v_l, v_r, a_l, a_r, g);
}
static star_values trrs_star(st5 l, st5 r, double v_l, double v_r, double a_l, double a_r, const adiabatic* g) {
  double beta = g->gm1d2g;
  double gm1d2 = 0.5 * (g->gamma - 1.);
  double num = a_l + a_r - gm1d2 * (v_r - v_l);
  double denom = a_l * pow(l.c[4], -beta) + a_r * pow(r.c[4], -beta);
  double pstar = pow(num / denom, 1. / beta);
  return star_state_from_pstar(pstar, l, r, v_l, v_r, a_l, a_r, g);
}

static int star_for_kind(int kind, double threshold, st5 l, st5 r, double v_l, double v_r, double a_l, double a_r,
                         const adiabatic* g, star_values* out, ora_counters* cnt) {
  switch (kind) {
    case ORA_EXACT:
      return exact_star(l, r, v_l, v_r, a_l, a_r, g, out, cnt);
    case ORA_PVRS:
      *out = pvrs_star(l, r, v_l, v_r, a_l, a_r);
      return ORA_OK;
    case ORA_AIRS: { /* airs.rs:19-37 */
      star_values star = pvrs_star(l, r, v_l, v_r, a_l, a_r);
      double p_max = rmax(l.c[4], r.c[4]);
      double p_min = rmin(l.c[4], r.c[4]);
      if (p_max / p_min < threshold) {
        *out = star;
        return ORA_OK;
      }
      return exact_star(l, r, v_l, v_r, a_l, a_r, g, out, cnt);
    }
    case ORA_TSRS:
      *out = tsrs_star(l, r, v_l, v_r, a_l, a_r, g);
      return ORA_OK;
    case ORA_TRRS:
      *out = trrs_star(l, r, v_l, v_r, a_l, a_r, g);
      return ORA_OK;
  }
  return ORA_ERR_UNSUPPORTED;
}

/* ---- HLLC (src/riemann_solver/hllc.rs) ---- */
static st5 hllc_flux(st5 s, double v, double e, vec3 n) {
  double rho_v = s.c[0] * v;
  return st_make(rho_v, v3_add(v3_scale(rho_v, st_vec(s)), v3_scale(s.c[4], n)), ((s.c[0] * e) + s.c[4]) * v);
}
static st5 hllc_flux_star(st5 s, double v, double e, double S, double s_star, vec3 n) {
  double s_m_v = S - v;
  double rho = s.c[0];
  double starfac = rho * s_m_v / (S - s_star);
  vec3 v_star = v3_add(v3_scale(s_star - v, n), st_vec(s));
  double e_star = e + (s_star - v) * (s_star + s.c[4] / (rho * s_m_v));
  st5 u_star = st_scale(starfac, st_make(1., v_star, e_star));
  st5 u_l = st_make(rho, v3_scale(rho, st_vec(s)), rho * e);
  return st_scale(S, st_sub(u_star, u_l));
}
static st5 hllc_solve(st5 left, st5 right, vec3 iv, vec3 n, const gaslaw* eos) {
  const adiabatic* g = &eos->gamma;
  double rho_l_inv = 1. / left.c[0];
  double rho_r_inv = 1. / right.c[0];
  double v_l = v3_dot(st_vec(left), n);
  double v_r = v3_dot(st_vec(right), n);
  double a_l = gas_sound_speed(eos, left.c[4], rho_l_inv);
  double a_r = gas_sound_speed(eos, right.c[4], rho_r_inv);
  double v_r_m_v_l = v_r - v_l;
  if (is_vacuum(left, right, a_l, a_r, v_r_m_v_l, g)) {
    st5 w_half = vacuum_sample(left, right, v_l, v_r, a_l, a_r, n, g);
    return flux_from_half_state(w_half, iv, n, g);
  }
  double ppvrs = 0.5 * (left.c[4] + right.c[4]) - 0.125 * v_r_m_v_l * (left.c[0] + right.c[0]) * (a_l + a_r);
  double p_star = rmax(ppvrs, 0.);
  double q_l = 1.;
  if (p_star > left.c[4] && left.c[4] > 0.) q_l = sqrt(1. + 0.5 * g->gp1dg * (p_star / left.c[4] - 1.));
  double q_r = 1.;
  if (p_star > right.c[4] && right.c[4] > 0.) q_r = sqrt(1. + 0.5 * g->gp1dg * (p_star / right.c[4] - 1.));
  double s_l_m_v_l = -a_l * q_l;
  double s_r_m_v_r = a_r * q_r;
  double s_star = (right.c[4] - left.c[4] + left.c[0] * v_l * s_l_m_v_l - right.c[0] * v_r * s_r_m_v_r) /
                  (left.c[0] * s_l_m_v_l - right.c[0] * s_r_m_v_r);
  st5 flux;
  if (s_star >= 0.) {
    double v_l2 = v3_len2(st_vec(left));
    double e_l = gas_internal_energy_from_pressure(eos, left.c[4], rho_l_inv) + 0.5 * v_l2;
    flux = hllc_flux(left, v_l, e_l, n);
    double s_l = s_l_m_v_l + v_l;
    if (s_l < 0.) flux = st_add(flux, hllc_flux_star(left, v_l, e_l, s_l, s_star, n));
  } else {
    double v_r2 = v3_len2(st_vec(right));
    double e_r = gas_internal_energy_from_pressure(eos, right.c[4], rho_r_inv) + 0.5 * v_r2;
    flux = hllc_flux(right, v_r, e_r, n);
    double s_r = s_r_m_v_r + v_r;
    if (s_r > 0.) flux = st_add(flux, hllc_flux_star(right, v_r, e_r, s_r, s_star, n));
  }
  /* deboost (hllc.rs:85-90) */
  flux = st_add(flux, st_make(0., v3_scale(flux.c[0], iv),
                              v3_dot(iv, st_vec(flux)) + 0.5 * v3_len2(iv) * flux.c[0]));
  return flux;
}

/* Blanket RiemannFluxSolver impl for star solvers (src/riemann_solver.rs:258-287) */
static int star_solver_flux(int kind, double threshold, st5 left, st5 right, vec3 iv, vec3 n, const gaslaw* eos,
                            st5* out, ora_counters* cnt) {
  const adiabatic* g = &eos->gamma;
  double v_l = v3_dot(st_vec(left), n);
  double v_r = v3_dot(st_vec(right), n);
  double a_l = gas_sound_speed(eos, left.c[4], 1. / left.c[0]);
  double a_r = gas_sound_speed(eos, right.c[4], 1. / right.c[0]);
  double v_r_m_v_l = v_r - v_l;
  st5 w_half;
  if (is_vacuum(left, right, a_l, a_r, v_r_m_v_l, g)) {
    w_half = vacuum_sample(left, right, v_l, v_r, a_l, a_r, n, g);
  } else {
    star_values star;
    int e = star_for_kind(kind, threshold, left, right, v_l, v_r, a_l, a_r, g, &star, cnt);
    if (e) return e;
    w_half = sample_from_star(star, left, right, v_l, v_r, a_l, a_r, n, g);
  }
  *out = flux_from_half_state(w_half, iv, n, g);
  return ORA_OK;
}

int ora_riemann_solve(const ora_riemann* rs, st5 left, st5 right, vec3 iv, vec3 n, const gaslaw* eos, st5* flux,
                      ora_counters* cnt) {
  switch (rs->kind) {
    case ORA_HLLC:
      *flux = hllc_solve(left, right, iv, n, eos);
      return ORA_OK;
    case ORA_LINADV: { /* linear_advection.rs:17-33 */
      vec3 vel = v3(rs->velocity[0], rs->velocity[1], rs->velocity[2]);
      double v = v3_dot(v3_sub(vel, iv), n);
      st5 s = v > 0. ? left : right;
      *flux = st_scale(v, s);
      return ORA_OK;
    }
    default:
      return star_solver_flux(rs->kind, rs->threshold, left, right, iv, n, eos, flux, cnt);
  }
}

/* ---- pairwise limiter (src/gradients.rs:105-152) ---- */
static double pairwise_single(double q_l, double q_r, double q_dash, double dx_fac) {
  if (q_l == q_r) return q_l;
  double q_bar = q_l + dx_fac * (q_r - q_l);
  double q_diff = fabs(q_l - q_r);
  double delta1 = 0.5 * q_diff;
  double delta2 = 0.25 * q_diff;
  if (q_l < q_r) {
    double q_min = rmin(q_l, q_r);
    double qmin = ((q_min - delta1) * q_min > 0.) ? q_min - delta1 : q_min * fabs(q_min) / (fabs(q_min) + delta1);
    return rmax(qmin, rmin(q_bar + delta2, q_dash));
  }
  double q_max = rmax(q_l, q_r);
  double qplu = ((q_max + delta1) * q_max > 0.) ? q_max + delta1 : q_max * fabs(q_max) / (fabs(q_max) + delta1);
  return rmin(qplu, rmax(q_bar - delta2, q_dash));
}
st5 ora_pairwise(st5 pl, st5 pr, st5 pdash, double dx_fac, ora_counters* cnt) {
  st5 lim = st_vacuum();
  for (int i = 0; i < 5; ++i) lim.c[i] = pairwise_single(pl.c[i], pr.c[i], pdash.c[i], dx_fac);
  ora_check_physical(&lim, cnt);
  return lim;
}

/* ---- exported unit API ---- */
static gaslaw ideal(double gamma) {
  gaslaw g;
  g.gamma = adiabatic_new(gamma);
  g.eos = ORA_IDEAL;
  g.iso_u = 0.;
  return g;
}
static st5 st_load(const double* p) { st5 s = {{p[0], p[1], p[2], p[3], p[4]}}; return s; }

int ora_solve_for_flux(const ora_riemann* rs, double gamma, const double* left, const double* right,
                       const double* iface_v, const double* n_unit, double* flux_out) {
  gaslaw g = ideal(gamma);
  st5 f;
  int e = ora_riemann_solve(rs, st_load(left), st_load(right), v3(iface_v[0], iface_v[1], iface_v[2]),
                            v3(n_unit[0], n_unit[1], n_unit[2]), &g, &f, 0);
  for (int i = 0; i < 5; ++i) flux_out[i] = f.c[i];
  return e;
}

int ora_sample(int kind, double threshold, double gamma, const double* left, const double* right, double v_l,
               double v_r, double a_l, double a_r, const double* n_unit, double* half_out) {
  adiabatic g = adiabatic_new(gamma);
  st5 l = st_load(left), r = st_load(right), h;
  vec3 n = v3(n_unit[0], n_unit[1], n_unit[2]);
  if (kind == 100) {
    h = vacuum_sample(l, r, v_l, v_r, a_l, a_r, n, &g);
  } else {
    star_values star;
    int e = star_for_kind(kind, threshold, l, r, v_l, v_r, a_l, a_r, &g, &star, 0);
    if (e) return e;
    h = sample_from_star(star, l, r, v_l, v_r, a_l, a_r, n, &g);
  }
  for (int i = 0; i < 5; ++i) half_out[i] = h.c[i];
  return ORA_OK;
}

int ora_exact_star(double gamma, const double* left, const double* right, double v_l, double v_r, double a_l,
                   double a_r, double* star_out) {
  adiabatic g = adiabatic_new(gamma);
  star_values sv;
  int e = exact_star(st_load(left), st_load(right), v_l, v_r, a_l, a_r, &g, &sv, 0);
  star_out[0] = sv.rho_l;
  star_out[1] = sv.rho_r;
  star_out[2] = sv.u;
  star_out[3] = sv.p;
  return e;
}

void ora_pairwise_limiter(const double* pl, const double* pr, const double* pdash, double dx_fac, double* out) {
  st5 r = ora_pairwise(st_load(pl), st_load(pr), st_load(pdash), dx_fac, 0);
  for (int i = 0; i < 5; ++i) out[i] = r.c[i];
}

double ora_sound_speed(double gamma, double pressure, double density_inv) {
  gaslaw g = ideal(gamma);
  return gas_sound_speed(&g, pressure, density_inv);
}
