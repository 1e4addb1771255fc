// Device physics of the per-face finite-volume update (gfx950).
//
// Templated on the arithmetic type T (double for the FP64 path, float for the FP32-flux path).  For
// T = double every expression keeps the reference's evaluation order (Rust evaluates left to right and
// never contracts to FMA; this file is compiled with -ffp-contract=off), so the FP64 kernels reproduce
// the reference bit for bit wherever +, -, *, /, sqrt are involved; pow goes through the device libm.
//
// Restates: src/physical_quantities.rs, src/gas_law.rs, src/riemann_solver.rs,
// src/riemann_solver/{exact,hllc,pvrs,airs,tsrs,trrs,vacuum,linear_advection}.rs,
// src/gradients.rs:105-152, src/flux.rs, src/part.rs:134-149.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfv {

// ------------------------------------------------------------------ error/counter codes
enum : int {
  E_NONFINITE = -1,
  E_NEG_MASS = -2,
  E_NEG_ENERGY = -3,
  E_NEG_PRESSURE = -4,
  E_NEWTON = -5,
  E_BRENT = -6,
  E_UNSUPPORTED = -7,
  E_DRIFT = -8,
  E_MESH = -22,  // the device tessellation failed (HFV_ERR_MESH)
};
enum { SOLVER_EXACT = 0, SOLVER_HLLC, SOLVER_PVRS, SOLVER_AIRS, SOLVER_TSRS, SOLVER_TRRS, SOLVER_LINADV };
enum { BND_PERIODIC = 0, BND_REFLECTIVE, BND_OPEN, BND_VACUUM };
enum { MOT_FIXED = 0, MOT_STEER, MOT_STEER_PAKMOR, MOT_FLUID };

// Per-thread soft-event counters and first hard error.
struct Events {
  int err;
  int check_physical;
  int brent;
  int grad_fallback;
  int flux_reset;
  int deferred;  // the face needs the exact solver and was left to the fallback pass
  __device__ __host__ Events() : err(0), check_physical(0), brent(0), grad_fallback(0), flux_reset(0), deferred(0) {}
  __device__ __host__ void fail(int e) {
    if (!err) err = e;
  }
};

// ------------------------------------------------------------------ math wrappers
__device__ __host__ __forceinline__ double hsqrt(double x) { return sqrt(x); }
__device__ __host__ __forceinline__ float hsqrt(float x) { return sqrtf(x); }
// Division and square root in the flux path's working type T: IEEE in FP64 (bit parity with the FP64
// reference); in the FP32 flux path the hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32,
// about 1 ulp each; two instructions instead of an IEEE-exact expansion of about ten), whose error is
// far inside the FP32 path's accuracy budget (L1 density error against FP64).
__device__ __host__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __host__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __host__ __forceinline__ double hpow(double x, double y) { return pow(x, y); }
__device__ __host__ __forceinline__ float hpow(float x, float y) { return powf(x, y); }
__device__ __host__ __forceinline__ double habs(double x) { return fabs(x); }
__device__ __host__ __forceinline__ float habs(float x) { return fabsf(x); }
// Rust f64::min / f64::max ignore a NaN operand (IEEE minNum / maxNum), like fmin / fmax.
__device__ __host__ __forceinline__ double hmin(double a, double b) { return fmin(a, b); }
__device__ __host__ __forceinline__ double hmax(double a, double b) { return fmax(a, b); }
__device__ __host__ __forceinline__ float hmin(float a, float b) { return fminf(a, b); }
__device__ __host__ __forceinline__ float hmax(float a, float b) { return fmaxf(a, b); }
__device__ __host__ __forceinline__ bool hfinite(double x) { return isfinite(x); }
__device__ __host__ __forceinline__ bool hfinite(float x) { return isfinite(x); }

// ------------------------------------------------------------------ glam::DVec3 semantics
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
__device__ __host__ __forceinline__ V3<T> mk3(T x, T y, T z) {
  V3<T> r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
template <class T>
__device__ __host__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) {
  return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class T>
__device__ __host__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) {
  return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class T>
__device__ __host__ __forceinline__ V3<T> operator-(V3<T> a) {
  return mk3(-a.x, -a.y, -a.z);
}
// f64 * DVec3
template <class T>
__device__ __host__ __forceinline__ V3<T> operator*(T s, V3<T> a) {
  return mk3(s * a.x, s * a.y, s * a.z);
}
// (x*x' + y*y') + z*z'
template <class T>
__device__ __host__ __forceinline__ T dot(V3<T> a, V3<T> b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <class T>
__device__ __host__ __forceinline__ T len2(V3<T> a) {
  return dot(a, a);
}
template <class T>
__device__ __host__ __forceinline__ T len(V3<T> a) {
  return fsqrt(dot(a, a));
}
template <class T>
__device__ __host__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
  return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
template <class T>
__device__ __host__ __forceinline__ bool finite3(V3<T> a) {
  return hfinite(a.x) && hfinite(a.y) && hfinite(a.z);
}
template <class T>
__device__ __host__ __forceinline__ T get3(V3<T> a, int i) {
  return i == 0 ? a.x : (i == 1 ? a.y : a.z);
}
template <class T, class U>
__device__ __host__ __forceinline__ V3<T> cvt3(V3<U> a) {
  return mk3((T)a.x, (T)a.y, (T)a.z);
}

// State<T>: c[0] = rho | mass, c[1..3] = v | momentum, c[4] = P | E
template <class T>
struct St {
  T c[5];
};
template <class T>
__device__ __host__ __forceinline__ St<T> mkst(T a, V3<T> v, T b) {
  St<T> s;
  s.c[0] = a;
  s.c[1] = v.x;
  s.c[2] = v.y;
  s.c[3] = v.z;
  s.c[4] = b;
  return s;
}
template <class T>
__device__ __host__ __forceinline__ V3<T> vel(const St<T>& s) {
  return mk3(s.c[1], s.c[2], s.c[3]);
}
template <class T>
__device__ __host__ __forceinline__ St<T> operator+(const St<T>& a, const St<T>& b) {
  St<T> r;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.c[i] = a.c[i] + b.c[i];
  return r;
}
template <class T>
__device__ __host__ __forceinline__ St<T> operator-(const St<T>& a, const St<T>& b) {
  St<T> r;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.c[i] = a.c[i] - b.c[i];
  return r;
}
template <class T>
__device__ __host__ __forceinline__ St<T> operator*(T s, const St<T>& a) {
  St<T> r;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.c[i] = s * a.c[i];
  return r;
}
template <class T>
__device__ __host__ __forceinline__ St<T> vacuum() {
  St<T> r;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.c[i] = T(0);
  return r;
}
template <class T, class U>
__device__ __host__ __forceinline__ St<T> cvtst(const St<U>& a) {
  St<T> r;
#pragma unroll
  for (int i = 0; i < 5; ++i) r.c[i] = (T)a.c[i];
  return r;
}

// Gradients<T>: g[q] = gradient of component q
template <class T>
struct Gr {
  V3<T> g[5];
};
template <class T>
__device__ __host__ __forceinline__ St<T> gdot(const Gr<T>& g, V3<T> dx) {
  return mkst(dot(g.g[0], dx), mk3(dot(g.g[1], dx), dot(g.g[2], dx), dot(g.g[3], dx)), dot(g.g[4], dx));
}
template <class T>
__device__ __host__ __forceinline__ T div_v(const Gr<T>& g) {
  return g.g[1].x + g.g[2].y + g.g[3].z;
}

// ------------------------------------------------------------------ gas law (src/gas_law.rs)
// AdiabaticIndex constants, computed on the host in double with the reference's expressions.
template <class T>
struct Gas {
  T gamma, gamma_inv, odgm1, odgp1, gp1dg, gm1d2g, gm1dgp1, tdgm1, tdgp1;
  T gamma_tdgm1;   // gamma * tdgm1  (riemann_solver.rs:132, exact.rs:92)
  T half_gm1;      // 0.5 * (gamma - 1)
  T one_over_gamma;// 1. / gamma (exact.rs:202)
  T iso_u;
  int isothermal;
};

template <class T>
__device__ __forceinline__ T gas_u_from_p(const Gas<T>& g, T p, T rho_inv) {
  if (!g.isothermal) return p * rho_inv * g.odgm1;
  return g.iso_u;
}
template <class T>
__device__ __forceinline__ T gas_p_from_u(const Gas<T>& g, T u, T rho) {
  if (!g.isothermal) return (g.gamma - T(1)) * u * rho;
  return (g.gamma - T(1)) * g.iso_u * rho;
}
template <class T>
__device__ __forceinline__ T sound_speed(const Gas<T>& g, T p, T rho_inv) {
  if (!g.isothermal) return fsqrt(g.gamma * p * rho_inv);
  return fsqrt(g.iso_u * g.gamma * (g.gamma - T(1)));
}

// ------------------------------------------------------------------ physical_quantities.rs:245-267
template <class T>
__device__ __forceinline__ void check_physical(St<T>& s, Events& ev) {
  if (s.c[0] < T(0)) {
    ev.check_physical++;
    s = vacuum<T>();
  }
  if (s.c[4] < T(0)) {
    ev.check_physical++;
    s = vacuum<T>();
  }
}
// physical_quantities.rs:228-236 (boost by `velocity`, a no-op for vacuum)
template <class T>
__device__ __forceinline__ St<T> boost(const St<T>& s, V3<T> velocity) {
  if (s.c[0] > T(0)) return mkst(s.c[0], vel(s) + velocity, s.c[4]);
  return s;
}
// physical_quantities.rs:239-242
template <class T>
__device__ __forceinline__ St<T> reflect(const St<T>& s, V3<T> n) {
  V3<T> v = vel(s);
  return mkst(s.c[0], v - ((T(2) * dot(v, n)) * n), s.c[4]);
}

// ------------------------------------------------------------------ riemann_solver.rs:46-76
template <class T>
__device__ __forceinline__ St<T> flux_from_half_state(const St<T>& half, V3<T> iv, V3<T> n, const Gas<T>& g) {
  V3<T> hv = vel(half);
  T rho = half.c[0], P = half.c[4];
  V3<T> v_tot = iv + hv;
  V3<T> f0 = rho * hv;
  V3<T> f1 = (rho * v_tot.x) * hv;
  f1.x += P;
  V3<T> f2 = (rho * v_tot.y) * hv;
  f2.y += P;
  V3<T> f3 = (rho * v_tot.z) * hv;
  f3.z += P;
  T roe = P * g.odgm1 + T(0.5) * rho * len2(v_tot);
  V3<T> f4 = (roe * hv) + (P * v_tot);
  return mkst(dot(f0, n), mk3(dot(f1, n), dot(f2, n), dot(f3, n)), dot(f4, n));
}

template <class T>
struct Star {
  T rho_l, rho_r, u, p;
};

// ------------------------------------------------------------------ riemann_solver.rs:96-256
template <class T>
__device__ __forceinline__ T shock_speed(T v, T a, T pdps, const Gas<T>& g) {
  return v - a * fsqrt(T(0.5) * g.gp1dg * pdps + g.gm1d2g);
}
template <class T>
__device__ __forceinline__ T rarefaction_tail_speed(T u, T a, T pdps, const Gas<T>& g) {
  return u - a * hpow(pdps, g.gm1d2g);
}
// rarefaction_tail_speed where only its sign against 0 is used.  FP64: the tail speed is first
// estimated with an FP32 pow; for pdps in [1e-30, 1e30] the estimate's relative error on a*pdps^e is
// below 2e-6 (a few ulp of powf, plus the rounding of pdps and e to FP32 amplified by |e ln pdps| < 35),
// so an estimate farther than 4e-5 |a y| (+ 1e-14 of the operands for the FP64 rounding) from 0 has the
// sign of the FP64 expression; only estimates inside that band use the FP64 pow.  Same decision as
// evaluating rarefaction_tail_speed in FP64, at a fraction of the cost on most faces.
template <class T>
__device__ __forceinline__ T tail_speed_for_sign(T u, T a, T pdps, const Gas<T>& g) {
  if constexpr (sizeof(T) == 8) {
    if (pdps > 1e-30 && pdps < 1e30) {
      const double y = (double)powf((float)pdps, (float)g.gm1d2g);
      const double ay = a * y;
      const double t = u - ay;
      const double bound = 4e-5 * fabs(ay) + 1e-14 * (fabs(u) + fabs(ay));
      if (t < -bound || t > bound) return t;
    }
  }
  return rarefaction_tail_speed(u, a, pdps, g);
}
template <class T>
__device__ __forceinline__ St<T> sample_rarefaction_fan(const St<T>& s, T a, T v, V3<T> n, const Gas<T>& g) {
  T v_half = g.tdgp1 * (a + g.half_gm1 * v);
  T base = g.tdgp1 + fdiv(g.gm1dgp1, a) * v;
  return mkst(s.c[0] * hpow(base, g.tdgm1), vel(s) + ((v_half - v) * n), s.c[4] * hpow(base, g.gamma_tdgm1));
}
template <class T>
__device__ __forceinline__ St<T> sample_middle_state(T rho, T u, T p, const St<T>& s, T v, V3<T> n) {
  return mkst(rho, vel(s) + ((u - v) * n), p);
}
template <class T>
__device__ __forceinline__ St<T> sample_from_star(const Star<T>& star, const St<T>& left, const St<T>& right, T v_l,
                                               T v_r, T a_l, T a_r, V3<T> n, const Gas<T>& g) {
  if (star.u < T(0)) {
    if (star.p > right.c[4]) {  // right shock
      if (shock_speed(v_r, -a_r, fdiv(star.p, right.c[4]), g) > T(0))
        return sample_middle_state(star.rho_r, star.u, star.p, right, v_r, n);
      return right;
    }
    // right rarefaction
    if (v_r - (-a_r) > T(0)) {
      if (tail_speed_for_sign(star.u, -a_r, fdiv(star.p, right.c[4]), g) < T(0))
        return sample_rarefaction_fan(right, -a_r, v_r, n, g);
      return sample_middle_state(star.rho_r, star.u, star.p, right, v_r, n);
    }
    return right;
  } else if (star.p > left.c[4]) {  // left shock
    if (shock_speed(v_l, a_l, fdiv(star.p, left.c[4]), g) < T(0))
      return sample_middle_state(star.rho_l, star.u, star.p, left, v_l, n);
    return left;
  }
  // left rarefaction
  if (v_l - a_l < T(0)) {
    if (tail_speed_for_sign(star.u, a_l, fdiv(star.p, left.c[4]), g) > T(0))
      return sample_rarefaction_fan(left, a_l, v_l, n, g);
    return sample_middle_state(star.rho_l, star.u, star.p, left, v_l, n);
  }
  return left;
}

// ------------------------------------------------------------------ vacuum.rs
template <class T>
__device__ __forceinline__ bool is_vacuum(const St<T>& left, const St<T>& right, T a_l, T a_r, T v_r_m_v_l,
                                          const Gas<T>& g) {
  if (left.c[0] == T(0) || right.c[0] == T(0)) return true;
  if (g.tdgm1 * (a_l + a_r) <= v_r_m_v_l) return true;
  return false;
}
template <class T>
// Out of line (rare path); states by value so that a call passes them in registers, not via scratch.
__device__ __noinline__ St<T> vacuum_sample(St<T> left, St<T> right, T v_l, T v_r, T a_l, T a_r, V3<T> n,
                                            const Gas<T>& g) {
  bool right_vacuum, left_vacuum;
  T s_l = T(0), s_r = T(0);
  if (right.c[0] == T(0)) {
    s_l = v_l + g.tdgm1 * a_l;
    right_vacuum = true;
    left_vacuum = false;
  } else if (left.c[0] == T(0)) {
    s_r = v_r - g.tdgm1 * a_r;
    right_vacuum = false;
    left_vacuum = true;
  } else {  // vacuum creation (vacuum.rs:78-99)
    s_l = v_l + g.tdgm1 * a_l;
    s_r = v_r - g.tdgm1 * a_r;
    if (T(0) <= s_l) {
      right_vacuum = true;
      left_vacuum = false;
    } else if (T(0) < s_r) {
      return vacuum<T>();
    } else {
      right_vacuum = false;
      left_vacuum = true;
    }
  }
  if (right_vacuum) {  // vacuum.rs:40-57
    if (T(0) <= v_l - a_l) return left;
    if (T(0) < s_l) return sample_rarefaction_fan(left, a_l, v_l, n, g);
    return vacuum<T>();
  }
  (void)left_vacuum;  // vacuum.rs:59-76
  if (T(0) <= s_r) return vacuum<T>();
  if (T(0) < v_r + a_r) return sample_rarefaction_fan(right, -a_r, v_r, n, g);
  return right;
}

// ------------------------------------------------------------------ exact.rs
template <class T>
__device__ __forceinline__ T ex_fb(T p, const St<T>& s, T a, const Gas<T>& g) {
  if (p > s.c[4]) {
    T cap_a = g.tdgp1 / s.c[0];
    T cap_b = g.gm1dgp1 * s.c[4];
    return (p - s.c[4]) * hsqrt(cap_a / (p + cap_b));
  }
  return g.tdgm1 * a * (hpow(p / s.c[4], g.gm1d2g) - T(1));
}
template <class T>
__device__ __forceinline__ T ex_f(T p, const St<T>& l, const St<T>& r, T v_l, T v_r, T a_l, T a_r, const Gas<T>& g) {
  return ex_fb(p, l, a_l, g) + ex_fb(p, r, a_r, g) + (v_r - v_l);
}
// exact.rs:37-45 (rarefaction branch exponent -(gamma-1)/(2 gamma), as in the reference)
template <class T>
__device__ __forceinline__ T ex_fprimeb(T p, const St<T>& s, T a, const Gas<T>& g) {
  if (p > s.c[4]) {
    T cap_a = g.tdgp1 / s.c[0];
    T cap_b = g.gm1dgp1 * s.c[4];
    return (T(1) - T(0.5) * (p - s.c[4]) / (cap_b + p)) * hsqrt(cap_a / (p + cap_b));
  }
  return T(1) / s.c[0] / a * hpow(p / s.c[4], -g.gm1d2g);
}
template <class T>
__device__ __forceinline__ T ex_gb(T p, const St<T>& s, const Gas<T>& g) {
  T cap_a = g.tdgp1 / s.c[0];
  T cap_b = g.gm1dgp1 * s.c[4];
  return hsqrt(cap_a / (p + cap_b));
}
template <class T>
__device__ __forceinline__ T ex_guess_p(const St<T>& l, const St<T>& r, T v_l, T v_r, T a_l, T a_r, const Gas<T>& g) {
  T p_min = hmin(l.c[4], r.c[4]);
  T p_max = hmax(l.c[4], r.c[4]);
  T q_max = p_max / p_min;
  T ppv = T(0.5) * (l.c[4] + r.c[4]) - T(0.125) * (v_r - v_l) * (l.c[0] + r.c[0]) * (a_l + a_r);
  ppv = hmax(ppv, T(1e-8));
  T p_guess;
  if (q_max <= T(2) && p_min <= ppv && ppv <= p_max) {
    p_guess = ppv;
  } else if (ppv < p_min) {
    T base = (a_l + a_r - g.half_gm1 * (v_r - v_l)) / (a_l / hpow(l.c[4], g.gm1d2g) + a_r / hpow(r.c[4], g.gm1d2g));
    p_guess = hpow(base, g.gamma_tdgm1);
  } else {
    T gl = ex_gb(ppv, l, g), gr = ex_gb(ppv, r, g);
    p_guess = (gl * l.c[4] + gr * r.c[4] - v_r + v_l) / (gl + gr);
  }
  return hmax(p_guess, T(1e-8));
}
template <class T>
__device__ __noinline__ T ex_solve_brent(T lower, T upper, T low_f, T up_f, T tol, St<T> l, St<T> r, T v_l, T v_r,
                                         T a_l, T a_r, const Gas<T>& g, Events& ev) {
  T a = lower, b = upper, c, d = T(INFINITY);
  T fa = low_f, fb = up_f, fc;
  T s, fs;
  if (fa * fb > T(0)) {
    ev.fail(E_BRENT);
    return b;
  }
  if (habs(fa) < habs(fb)) {
    T t = a; a = b; b = t;
    t = fa; fa = fb; fb = t;
  }
  c = a;
  fc = fa;
  bool mflag = true;
  int guard = 0;
  while (fb != T(0) && habs(a - b) > tol * T(0.5) * (a + b)) {
    if (fa != fc && fb != fc) {
      s = a * fb * fc / (fa - fb) / (fa - fc) + b * fa * fc / (fb - fa) / (fb - fc) + c * fa * fb / (fc - fa) / (fc - fb);
    } else {
      s = b - fb * (b - a) / (fb - fa);
    }
    T tmp = T(0.25) * (T(3) * a + b);
    if (!((s > tmp && s < b) || (s < tmp && s > b)) || (mflag && habs(s - b) >= (T(0.5) * habs(b - c))) ||
        (!mflag && habs(s - b) >= (T(0.5) * habs(c - d))) || (mflag && habs(b - c) < T(0.5) * tol * (b + c)) ||
        (!mflag && habs(c - d) < T(0.5) * tol * (c + d))) {
      s = T(0.5) * (a + b);
      mflag = true;
    } else {
      mflag = false;
    }
    fs = ex_f(s, l, r, v_l, v_r, a_l, a_r, g);
    d = c;
    c = b;
    fc = fb;
    if (fa * fs < T(0)) {
      b = s;
      fb = fs;
    } else {
      a = s;
      fa = fs;
    }
    if (habs(fa) < habs(fb)) {
      T t = a; a = b; b = t;
      t = fa; fa = fb; fb = t;
    }
    // bisection halves the bracket at least every other step; this bound is never reached for
    // finite inputs and only keeps a wave from spinning on NaN brackets.
    if (++guard > 100000) {
      ev.fail(E_BRENT);
      break;
    }
  }
  return b;
}
template <class T>
__device__ __forceinline__ T middle_density(T p, const St<T>& s, const Gas<T>& g) {
  T pdps = p / s.c[4];
  if (pdps > T(1)) return s.c[0] * (pdps + g.gm1dgp1) / (g.gm1dgp1 * pdps + T(1));
  return s.c[0] * hpow(pdps, g.one_over_gamma);
}
template <class T>
__device__ __forceinline__ Star<T> star_from_pstar(T pstar, const St<T>& l, const St<T>& r, T v_l, T v_r, T a_l,
                                                   T a_r, const Gas<T>& g) {
  Star<T> sv;
  sv.u = T(0.5) * (v_l + v_r) + T(0.5) * (ex_fb(pstar, r, a_r, g) - ex_fb(pstar, l, a_l, g));
  sv.rho_l = middle_density(pstar, l, g);
  sv.rho_r = middle_density(pstar, r, g);
  sv.p = pstar;
  return sv;
}
// exact.rs:242-290
template <class T>
__device__ __noinline__ Star<T> exact_star(St<T> l, St<T> r, T v_l, T v_r, T a_l, T a_r, const Gas<T>& g,
                                           Events& ev) {
  T p = T(0);
  T p_guess = ex_guess_p(l, r, v_l, v_r, a_l, a_r, g);
  T fp = ex_f(p, l, r, v_l, v_r, a_l, a_r, g);
  T fp_guess = ex_f(p_guess, l, r, v_l, v_r, a_l, a_r, g);
  if (fp * fp_guess >= T(0)) {
    int counter = 0;
    while (habs(p - p_guess) > T(1e-6) * T(0.5) * (p + p_guess) && fp_guess < T(0)) {
      p = p_guess;
      T fpr = ex_fprimeb(p_guess, l, a_l, g) + ex_fprimeb(p_guess, r, a_r, g);
      p_guess = p_guess - fp_guess / fpr;
      fp_guess = ex_f(p_guess, l, r, v_l, v_r, a_l, a_r, g);
      counter += 1;
      if (counter > 1000) {
        ev.fail(E_NEWTON);
        break;
      }
    }
  }
  if (habs(p - p_guess) > T(1e-6) * T(0.5) * (p + p_guess) && fp_guess > T(0)) {
    p = T(0);
    fp = ex_f(p, l, r, v_l, v_r, a_l, a_r, g);
    ev.brent++;
    p = ex_solve_brent(p, p_guess, fp, fp_guess, T(1e-6), l, r, v_l, v_r, a_l, a_r, g, ev);
  } else {
    p = p_guess;
  }
  return star_from_pstar(p, l, r, v_l, v_r, a_l, a_r, g);
}

// ------------------------------------------------------------------ pvrs.rs / tsrs.rs / trrs.rs
template <class T>
__device__ __forceinline__ T pv_p_star(T rho_bar, T p_bar, T a_bar, T v_l, T v_r) {
  return p_bar + T(0.5) * (v_l - v_r) * rho_bar * a_bar;
}
template <class T>
__device__ __forceinline__ Star<T> pvrs_star(const St<T>& l, const St<T>& r, T v_l, T v_r, T a_l, T a_r) {
  T rho_bar = T(0.5) * (l.c[0] + r.c[0]);
  T p_bar = T(0.5) * (l.c[4] + r.c[4]);
  T a_bar = T(0.5) * (a_l + a_r);
  Star<T> sv;
  sv.p = pv_p_star(rho_bar, p_bar, a_bar, v_l, v_r);
  sv.u = T(0.5) * ((v_l + v_r) + fdiv(l.c[4] - r.c[4], rho_bar * a_bar));
  sv.rho_l = l.c[0] + fdiv((v_l - sv.u) * rho_bar, a_bar);
  sv.rho_r = r.c[0] + fdiv((sv.u - v_r) * rho_bar, a_bar);
  return sv;
}
template <class T>
__device__ __forceinline__ Star<T> tsrs_star(const St<T>& l, const St<T>& r, T v_l, T v_r, T a_l, T a_r,
                                             const Gas<T>& g) {
  T p_guess = pv_p_star(T(0.5) * (l.c[0] + r.c[0]), T(0.5) * (l.c[4] + r.c[4]), T(0.5) * (a_l + a_r), v_l, v_r);
  T g_l = hsqrt((g.tdgp1 / l.c[0]) / (p_guess + g.gm1dgp1 * l.c[4]));
  T g_r = hsqrt((g.tdgp1 / r.c[0]) / (p_guess + g.gm1dgp1 * r.c[4]));
  T pstar = (g_l * l.c[4] + g_r * r.c[4] - (v_r - v_l)) / (g_l + g_r);
  return star_from_pstar(pstar, l, r, v_l, v_r, a_l, a_r, g);
}
template <class T>
__device__ __forceinline__ Star<T> trrs_star(const St<T>& l, const St<T>& r, T v_l, T v_r, T a_l, T a_r,
                                             const Gas<T>& g) {
  T beta = g.gm1d2g;
  T num = a_l + a_r - g.half_gm1 * (v_r - v_l);
  T denom = a_l * hpow(l.c[4], -beta) + a_r * hpow(r.c[4], -beta);
  T pstar = hpow(num / denom, T(1) / beta);
  return star_from_pstar(pstar, l, r, v_l, v_r, a_l, a_r, g);
}

// ------------------------------------------------------------------ hllc.rs
template <class T>
__device__ __forceinline__ St<T> hllc_flux(const St<T>& s, T v, T e, V3<T> n) {
  T rho_v = s.c[0] * v;
  return mkst(rho_v, (rho_v * vel(s)) + (s.c[4] * n), ((s.c[0] * e) + s.c[4]) * v);
}
template <class T>
__device__ __forceinline__ St<T> hllc_flux_star(const St<T>& s, T v, T e, T S, T s_star, V3<T> n) {
  T s_m_v = S - v;
  T rho = s.c[0];
  T starfac = fdiv(rho * s_m_v, S - s_star);
  V3<T> v_star = ((s_star - v) * n) + vel(s);
  T e_star = e + (s_star - v) * (s_star + fdiv(s.c[4], rho * s_m_v));
  St<T> u_star = starfac * mkst(T(1), v_star, e_star);
  St<T> u_l = mkst(rho, rho * vel(s), rho * e);
  return S * (u_star - u_l);
}

// ------------------------------------------------------------------ solver dispatch
struct SolverCfg {
  int kind;
  double threshold;
  double linv[3];
};

// FP64 star state for any star solver (exact/PVRS/AIRS/TSRS/TRRS).
template <class T>
__device__ __forceinline__ Star<T> star_for(int kind, double threshold, const St<T>& l, const St<T>& r, T v_l, T v_r,
                                            T a_l, T a_r, const Gas<T>& g, Events& ev) {
  switch (kind) {
    case SOLVER_PVRS:
      return pvrs_star(l, r, v_l, v_r, a_l, a_r);
    case SOLVER_AIRS: {  // airs.rs:19-37
      Star<T> st = pvrs_star(l, r, v_l, v_r, a_l, a_r);
      T p_max = hmax(l.c[4], r.c[4]);
      T p_min = hmin(l.c[4], r.c[4]);
      if (fdiv(p_max, p_min) < (T)threshold) return st;
      return exact_star(l, r, v_l, v_r, a_l, a_r, g, ev);
    }
    case SOLVER_TSRS:
      return tsrs_star(l, r, v_l, v_r, a_l, a_r, g);
    case SOLVER_TRRS:
      return trrs_star(l, r, v_l, v_r, a_l, a_r, g);
    default:
      return exact_star(l, r, v_l, v_r, a_l, a_r, g, ev);
  }
}

// RiemannFluxSolver::solve_for_flux.  For the FP32-flux path (T = float) the exact solver's pressure
// iteration runs in FP64 (gd = double constants) and its star state is rounded back to float.
// DEFER (AIRS only): a face whose pressure ratio calls for the exact solver is not solved here
// (ev.deferred is set); the caller hands it to a separate pass, so that this code path carries no
// exact-solver registers.
template <class T, bool DEFER = false>
__device__ __forceinline__ St<T> riemann_flux(const SolverCfg& sc, const St<T>& left, const St<T>& right, V3<T> iv, V3<T> n,
                              const Gas<T>& g, const Gas<double>& gd, Events& ev) {
  if (sc.kind == SOLVER_LINADV) {  // linear_advection.rs:17-33
    V3<T> velo = mk3((T)sc.linv[0], (T)sc.linv[1], (T)sc.linv[2]);
    T v = dot(velo - iv, n);
    return v * (v > T(0) ? left : right);
  }
  if (sc.kind == SOLVER_HLLC) {  // hllc.rs:12-92
    T rho_l_inv = fdiv(T(1), left.c[0]);
    T rho_r_inv = fdiv(T(1), right.c[0]);
    T v_l = dot(vel(left), n);
    T v_r = dot(vel(right), n);
    T a_l = sound_speed(g, left.c[4], rho_l_inv);
    T a_r = sound_speed(g, right.c[4], rho_r_inv);
    T v_r_m_v_l = v_r - v_l;
    if (is_vacuum(left, right, a_l, a_r, v_r_m_v_l, g)) {
      St<T> w = vacuum_sample(left, right, v_l, v_r, a_l, a_r, n, g);
      return flux_from_half_state(w, iv, n, g);
    }
    T ppvrs = T(0.5) * (left.c[4] + right.c[4]) - T(0.125) * v_r_m_v_l * (left.c[0] + right.c[0]) * (a_l + a_r);
    T p_star = hmax(ppvrs, T(0));
    T q_l = T(1);
    if (p_star > left.c[4] && left.c[4] > T(0)) q_l = fsqrt(T(1) + T(0.5) * g.gp1dg * (fdiv(p_star, left.c[4]) - T(1)));
    T q_r = T(1);
    if (p_star > right.c[4] && right.c[4] > T(0)) q_r = fsqrt(T(1) + T(0.5) * g.gp1dg * (fdiv(p_star, right.c[4]) - T(1)));
    T s_l_m_v_l = -a_l * q_l;
    T s_r_m_v_r = a_r * q_r;
    T s_star = fdiv(right.c[4] - left.c[4] + left.c[0] * v_l * s_l_m_v_l - right.c[0] * v_r * s_r_m_v_r,
                    left.c[0] * s_l_m_v_l - right.c[0] * s_r_m_v_r);
    St<T> flux;
    if (s_star >= T(0)) {
      T v_l2 = len2(vel(left));
      T e_l = gas_u_from_p(g, left.c[4], rho_l_inv) + T(0.5) * v_l2;
      flux = hllc_flux(left, v_l, e_l, n);
      T s_l = s_l_m_v_l + v_l;
      if (s_l < T(0)) flux = flux + hllc_flux_star(left, v_l, e_l, s_l, s_star, n);
    } else {
      T v_r2 = len2(vel(right));
      T e_r = gas_u_from_p(g, right.c[4], rho_r_inv) + T(0.5) * v_r2;
      flux = hllc_flux(right, v_r, e_r, n);
      T s_r = s_r_m_v_r + v_r;
      if (s_r > T(0)) flux = flux + hllc_flux_star(right, v_r, e_r, s_r, s_star, n);
    }
    flux = flux + mkst(T(0), flux.c[0] * iv, dot(iv, vel(flux)) + T(0.5) * len2(iv) * flux.c[0]);
    return flux;
  }
  // blanket impl for star solvers (riemann_solver.rs:258-287)
  T v_l = dot(vel(left), n);
  T v_r = dot(vel(right), n);
  T a_l = sound_speed(g, left.c[4], fdiv(T(1), left.c[0]));
  T a_r = sound_speed(g, right.c[4], fdiv(T(1), right.c[0]));
  T v_r_m_v_l = v_r - v_l;
  St<T> w;
  if (is_vacuum(left, right, a_l, a_r, v_r_m_v_l, g)) {
    w = vacuum_sample(left, right, v_l, v_r, a_l, a_r, n, g);
  } else if (DEFER) {
    if (!(fdiv(hmax(left.c[4], right.c[4]), hmin(left.c[4], right.c[4])) < (T)sc.threshold)) {
      ev.deferred = 1;
      return vacuum<T>();
    }
    w = sample_from_star(pvrs_star(left, right, v_l, v_r, a_l, a_r), left, right, v_l, v_r, a_l, a_r, n, g);
  } else {
    Star<T> star;
    bool need_fp64 = (sizeof(T) != sizeof(double)) &&
                     (sc.kind == SOLVER_EXACT ||
                      (sc.kind == SOLVER_AIRS && !(fdiv(hmax(left.c[4], right.c[4]), hmin(left.c[4], right.c[4])) <
                                                   (T)sc.threshold)));
    if (need_fp64) {
      St<double> ld = cvtst<double>(left), rd = cvtst<double>(right);
      Star<double> sd = exact_star(ld, rd, (double)v_l, (double)v_r, (double)a_l, (double)a_r, gd, ev);
      star.rho_l = (T)sd.rho_l;
      star.rho_r = (T)sd.rho_r;
      star.u = (T)sd.u;
      star.p = (T)sd.p;
    } else {
      star = star_for(sc.kind, sc.threshold, left, right, v_l, v_r, a_l, a_r, g, ev);
    }
    w = sample_from_star(star, left, right, v_l, v_r, a_l, a_r, n, g);
  }
  return flux_from_half_state(w, iv, n, g);
}

// ------------------------------------------------------------------ gradients.rs:105-152
template <class T>
__device__ __forceinline__ T pairwise_single(T q_l, T q_r, T q_dash, T dx_fac) {
  if (q_l == q_r) return q_l;
  T q_bar = q_l + dx_fac * (q_r - q_l);
  T q_diff = habs(q_l - q_r);
  T delta1 = T(0.5) * q_diff;
  T delta2 = T(0.25) * q_diff;
  if (q_l < q_r) {
    T q_min = hmin(q_l, q_r);
    T qmin = ((q_min - delta1) * q_min > T(0)) ? q_min - delta1 : fdiv(q_min * habs(q_min), habs(q_min) + delta1);
    return hmax(qmin, hmin(q_bar + delta2, q_dash));
  }
  T q_max = hmax(q_l, q_r);
  T qplu = ((q_max + delta1) * q_max > T(0)) ? q_max + delta1 : fdiv(q_max * habs(q_max), habs(q_max) + delta1);
  return hmin(qplu, hmax(q_bar - delta2, q_dash));
}
template <class T>
__device__ __forceinline__ St<T> pairwise_limiter(const St<T>& pl, const St<T>& pr, const St<T>& pdash, T dx_fac,
                                                  Events& ev) {
  St<T> lim;
#pragma unroll
  for (int i = 0; i < 5; ++i) lim.c[i] = pairwise_single(pl.c[i], pr.c[i], pdash.c[i], dx_fac);
  check_physical(lim, ev);
  return lim;
}

// ------------------------------------------------------------------ part.rs:134-149
// The state that time_extrapolations scales by -dt (valid for rho > 0).
template <class T>
__device__ __forceinline__ St<T> tex_rate(const St<T>& prim, const Gr<T>& gr, V3<T> v_rel, const Gas<T>& g) {
  T rho = prim.c[0];
  T rho_inv = T(1) / rho;
  T P = prim.c[4];
  T dv = div_v(gr);
  return mkst(rho * dv + dot(v_rel, gr.g[0]), (dv * v_rel) + (rho_inv * gr.g[4]),
              g.gamma * P * dv + dot(v_rel, gr.g[4]));
}
template <class T>
__device__ __forceinline__ St<T> time_extrapolations(const St<T>& prim, const Gr<T>& gr, V3<T> v_rel, T dt,
                                                     const Gas<T>& g) {
  if (prim.c[0] > T(0)) return (-dt) * tex_rate(prim, gr, v_rel, g);
  return vacuum<T>();
}

}  // namespace hfv
