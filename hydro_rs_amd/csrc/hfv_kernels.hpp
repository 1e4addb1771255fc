// Device kernels of the per-face finite-volume step (gfx950).  Included by hfv_api.hip.
//
// HBM layout (one context = one simulation):
//   cells, component-major SoA (field[c * N + i]): prim[5], grad[15], gc[3], extrap[5], cons[5],
//     facc[5] (pending flux accumulator), vol, loc[3], cen[3], v[3], vrel[3], maor, msv, dt, srad,
//     bin (int8), wake (int8), act (uint8)
//   faces, SoA: area, nrm[3], fcen[3], left (int32), right (int32, -1 = wall), shc (uint8 periodic
//     shift code, 2 bits per axis), rec[7] (dt*F[5], v_max, a_over_r)
//   csr: off[N+1] (int32), conn[2F] (int32: face << 1 | side, side 1 = the cell is the face's right)
#pragma once
#include "hfv_physics.hpp"

namespace hfv {

struct Cells {
  int64_t n;
  double *prim, *grad, *gc, *extrap, *cons, *facc, *vol, *loc, *cen, *v, *vrel, *maor, *msv, *dt, *srad;
  int8_t *bin, *wake;
  uint8_t* act;
};

struct Faces {
  int64_t nf;
  double *area, *nrm, *fcen, *rec;
  int32_t *left, *right, *off, *conn;
  uint8_t* shc;
};

struct Cfg {
  int dim, boundary, motion, runner, sync_all;
  double box[3];
  double cfl;
  double radius3_fac;  // 0.25 * 3. * FRAC_1_PI
  SolverCfg solver;
  Gas<double> gd;
  Gas<float> gf;
};

// Device-side counters of one context.
struct DevCounters {
  int err;
  int pad;
  unsigned long long check_physical, flux_reset, grad_fallback, brent;
  unsigned long long n_active, n_face_active;
  unsigned long long ti_min;
  int min_bin;
  int pad2;
};

#define HFV_NUM_TIME_BINS 56
#define HFV_MAX_NR_TIMESTEPS (1ull << 57)

// ------------------------------------------------------------------ timeline (src/timeline.rs)
__device__ __host__ __forceinline__ uint64_t get_integer_timestep(int bin) {
  if (bin <= 0) return 0;
  return 1ull << (bin + 1);
}
__device__ __host__ __forceinline__ int8_t get_time_bin(uint64_t dti) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t lz = dti ? (uint32_t)__clzll((long long)dti) : 64u;
#else
  uint32_t lz = dti ? (uint32_t)__builtin_clzll(dti) : 64u;
#endif
  uint32_t r = 64u - lz - 2u;
  return (int8_t)(uint8_t)(r & 0xFFu);
}
__device__ __host__ __forceinline__ uint64_t get_integer_time_end(uint64_t ti_current, int bin) {
  uint64_t dti = get_integer_timestep(bin);
  if (dti == 0) return 0;
  uint64_t m = ti_current % dti;
  if (m == 0) return ti_current;
  return ti_current - m + dti;
}
__device__ __host__ __forceinline__ int8_t get_max_active_bin(uint64_t ti) {
  if (ti == 0) return HFV_NUM_TIME_BINS;
  int bin = 1;
  while (bin < 62 && ((1ull << (bin + 1)) & ti) == 0) bin += 1;
  return (int8_t)bin;
}
__device__ __host__ __forceinline__ uint64_t sat_u64(double x) {
  if (!(x > 0.)) return 0;
  if (x >= 18446744073709551616.0) return ~0ull;
  return (uint64_t)x;
}
__device__ __host__ __forceinline__ uint64_t make_integer_timestep(double dt, int8_t old_bin, int8_t min_ngb_bin,
                                                                    uint64_t ti_current, double time_base_inv) {
  uint64_t new_dti = sat_u64(dt * time_base_inv);
  int8_t b = get_time_bin(new_dti);
  int8_t lim = (int8_t)(min_ngb_bin + 2);
  int8_t new_bin = b < lim ? b : lim;
  new_dti = get_integer_timestep(new_bin);
  uint64_t current_dti = get_integer_timestep(old_bin);
  if (old_bin > 0) {
    uint64_t t = 2 * current_dti;
    if (t < new_dti) new_dti = t;
  }
  uint64_t timeline_dti = HFV_MAX_NR_TIMESTEPS;
  while (new_dti < timeline_dti) timeline_dti /= 2;
  new_dti = timeline_dti;
  if (new_dti > current_dti) {
    uint64_t ti_end = get_integer_time_end(ti_current, old_bin);
    if ((HFV_MAX_NR_TIMESTEPS - ti_end) % new_dti != 0) new_dti = current_dti;
  }
  return new_dti;
}

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ V3<double> ld3(const double* a, int64_t n, int64_t i) {
  return mk3(a[i], a[n + i], a[2 * n + i]);
}
__device__ __forceinline__ void st3(double* a, int64_t n, int64_t i, V3<double> v) {
  a[i] = v.x;
  a[n + i] = v.y;
  a[2 * n + i] = v.z;
}
__device__ __forceinline__ St<double> ld5(const double* a, int64_t n, int64_t i) {
  St<double> s;
#pragma unroll
  for (int k = 0; k < 5; ++k) s.c[k] = a[k * n + i];
  return s;
}
__device__ __forceinline__ void st5(double* a, int64_t n, int64_t i, const St<double>& s) {
#pragma unroll
  for (int k = 0; k < 5; ++k) a[k * n + i] = s.c[k];
}
__device__ __forceinline__ Gr<double> ldg(const double* a, int64_t n, int64_t i) {
  Gr<double> g;
#pragma unroll
  for (int k = 0; k < 5; ++k) g.g[k] = mk3(a[(3 * k) * n + i], a[(3 * k + 1) * n + i], a[(3 * k + 2) * n + i]);
  return g;
}
__device__ __forceinline__ void stg(double* a, int64_t n, int64_t i, const Gr<double>& g) {
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    a[(3 * k) * n + i] = g.g[k].x;
    a[(3 * k + 1) * n + i] = g.g[k].y;
    a[(3 * k + 2) * n + i] = g.g[k].z;
  }
}
// periodic shift code: 2 bits per axis, 0 -> 0, 1 -> +1, 2 -> -1
__device__ __forceinline__ V3<double> decode_shift(uint8_t code, const double* box) {
  double s[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int c = (code >> (2 * a)) & 3;
    double m = c == 1 ? 1.0 : (c == 2 ? -1.0 : 0.0);
    s[a] = m * box[a];
  }
  return mk3(s[0], s[1], s[2]);
}

__device__ __forceinline__ void flush_events(const Events& ev, DevCounters* cnt) {
  if (ev.err) atomicCAS(&cnt->err, 0, ev.err);
  if (ev.check_physical) atomicAdd(&cnt->check_physical, (unsigned long long)ev.check_physical);
  if (ev.brent) atomicAdd(&cnt->brent, (unsigned long long)ev.brent);
  if (ev.grad_fallback) atomicAdd(&cnt->grad_fallback, (unsigned long long)ev.grad_fallback);
  if (ev.flux_reset) atomicAdd(&cnt->flux_reset, (unsigned long long)ev.flux_reset);
}

// Wavefront-aggregated count: one atomic per wave.
__device__ __forceinline__ void wave_count(bool pred, unsigned long long* dst) {
  unsigned long long m = __ballot(pred);
  if (__lane_id() == __ffsll((long long)__ballot(1)) - 1 && m) atomicAdd(dst, (unsigned long long)__popcll(m));
}

// ------------------------------------------------------------------ kernels
__global__ void k_reset_counters(DevCounters* cnt) {
  cnt->n_active = 0;
  cnt->n_face_active = 0;
  cnt->ti_min = ~0ull;
  cnt->min_bin = 127;
}

__global__ void k_clear_all_counters(DevCounters* cnt) {
  cnt->err = 0;
  cnt->check_physical = 0;
  cnt->flux_reset = 0;
  cnt->grad_fallback = 0;
  cnt->brent = 0;
  cnt->n_active = 0;
  cnt->n_face_active = 0;
  cnt->ti_min = ~0ull;
  cnt->min_bin = 127;
}

// Active mask (optimal_order.rs:17) fused with EulerEqnsFvs::predict (euler_eqns_fvs.rs:20-24,
// part.rs:123-149) over all cells.  do_predict = 0 computes the mask only.
__global__ void k_mask_predict(Cells C, Cfg cfg, double dt, int8_t mab, int do_predict, DevCounters* cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = false;
  if (i < C.n) {
    act = C.bin[i] <= mab;
    C.act[i] = act;
    if (do_predict) {
      Events ev;
      int64_t n = C.n;
      St<double> prim = ld5(C.prim, n, i);
      Gr<double> gr = ldg(C.grad, n, i);
      V3<double> vrel = ld3(C.vrel, n, i);
      St<double> ex = ld5(C.extrap, n, i);
      ex = ex + time_extrapolations(prim, gr, vrel, dt, cfg.gd);
      st5(C.extrap, n, i, ex);
      check_physical(prim, ev);
      if (ev.check_physical) {
        st5(C.prim, n, i, prim);
        flush_events(ev, cnt);
      }
      // volume += dvdt * dt clamped to [V/2, 2V]: dvdt is never set on the path (part.rs:130-131)
    }
  }
  wave_count(act, &cnt->n_active);
}

// FiniteVolumeSolver::compute_fluxes: gate (euler_eqns_fvs.rs:26-48) + flux_exchange (flux.rs:23-100)
// + flux_exchange_boundary (flux.rs:102-195).  T is the arithmetic of reconstruction and solve.
template <class T>
__device__ __forceinline__ const Gas<T>& gas_for(const Cfg& c);
template <>
__device__ __forceinline__ const Gas<double>& gas_for<double>(const Cfg& c) {
  return c.gd;
}
template <>
__device__ __forceinline__ const Gas<float>& gas_for<float>(const Cfg& c) {
  return c.gf;
}

template <class T>
__global__ void __launch_bounds__(256) k_flux(Cells C, Faces F, Cfg cfg, DevCounters* cnt) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool computed = false;
  if (f < F.nf) {
    const int64_t n = C.n;
    const int64_t nf = F.nf;
    int32_t l = F.left[f];
    int32_t r = F.right[f];
    bool la = C.act[l];
    double outv[7] = {0., 0., 0., 0., 0., 0., 0.};
    Events ev;
    const Gas<T>& g = gas_for<T>(cfg);
    if (r >= 0) {
      bool ra = C.act[r];
      double dtl = C.dt[l], dtr = C.dt[r];
      if (!((!la && !ra) || (dtr < dtl && !ra) || (dtl < dtr && !la))) {
        computed = true;
        double dt = hmin(dtl, dtr);
        V3<double> fc = ld3(F.fcen, nf, f);
        V3<double> nrm = ld3(F.nrm, nf, f);
        double area = F.area[f];
        V3<double> sh = decode_shift(F.shc[f], cfg.box);
        V3<double> gcl = ld3(C.gc, n, l), gcr = ld3(C.gc, n, r);
        V3<double> dx_left = fc - gcl;
        V3<double> dx_right = (fc - gcr) - sh;
        double dx_gradients = len(dx_left - dx_right);
        V3<double> cl = ld3(C.cen, n, l), cr = ld3(C.cen, n, r);
        V3<double> dx_centroid = (cr + sh) - cl;
        V3<double> locl = ld3(C.loc, n, l), locr = ld3(C.loc, n, r);
        V3<double> dx = (locr + sh) - locl;
        St<double> pl = ld5(C.prim, n, l), pr = ld5(C.prim, n, r);
        double v_max = 0.;
        if (pl.c[0] > 0.) v_max += sound_speed(cfg.gd, pl.c[4], 1. / pl.c[0]);
        if (pr.c[0] > 0.) v_max += sound_speed(cfg.gd, pr.c[4], 1. / pr.c[0]);
        v_max -= hmin(dot(vel(pr) - vel(pl), dx_centroid), 0.);
        // reconstruction (flux.rs:52-73) in T
        T dt_extrapolate = (T)(-0.5 * dt);
        Gr<double> gl = ldg(C.grad, n, l), grr = ldg(C.grad, n, r);
        St<double> exl = ld5(C.extrap, n, l), exr = ld5(C.extrap, n, r);
        V3<double> vrl = ld3(C.vrel, n, l), vrr = ld3(C.vrel, n, r);
        St<T> plT = cvtst<T>(pl), prT = cvtst<T>(pr);
        Gr<T> glT, grT;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          glT.g[q] = cvt3<T>(gl.g[q]);
          grT.g[q] = cvt3<T>(grr.g[q]);
        }
        V3<T> dxlT = cvt3<T>(dx_left), dxrT = cvt3<T>(dx_right);
        St<T> left_dash = ((plT + gdot(glT, dxlT)) + cvtst<T>(exl)) +
                          time_extrapolations(plT, glT, cvt3<T>(vrl), dt_extrapolate, g);
        St<T> right_dash = ((prT + gdot(grT, dxrT)) + cvtst<T>(exr)) +
                           time_extrapolations(prT, grT, cvt3<T>(vrr), dt_extrapolate, g);
        T dxg = (T)dx_gradients;
        St<T> prim_left = pairwise_limiter(plT, prT, left_dash, len(dxlT) / dxg, ev);
        St<T> prim_right = pairwise_limiter(prT, plT, right_dash, len(dxrT) / dxg, ev);
        // face velocity (flux.rs:75-78), FP64
        V3<double> vl = ld3(C.v, n, l), vr = ld3(C.v, n, r);
        V3<double> midpoint = 0.5 * ((locl + locr) + sh);
        double fac = dot(vr - vl, fc - midpoint) / len2(dx);
        V3<double> v_face = (0.5 * (vl + vr)) - (fac * dx);
        V3<T> vfT = cvt3<T>(v_face);
        St<T> Fl = riemann_flux(cfg.solver, boost(prim_left, -vfT), boost(prim_right, -vfT), vfT, cvt3<T>(nrm), g,
                                cfg.gd, ev);
        St<double> fluxes = area * cvtst<double>(Fl);
        St<double> rec = dt * fluxes;
#pragma unroll
        for (int k = 0; k < 5; ++k) outv[k] = rec.c[k];
        outv[5] = v_max;
        outv[6] = area / len(dx);
      }
    } else if (la) {
      computed = true;
      V3<double> fc = ld3(F.fcen, nf, f);
      V3<double> nrm = ld3(F.nrm, nf, f);
      double area = F.area[f];
      V3<double> loc = ld3(C.loc, n, l), cen = ld3(C.cen, n, l), gc = ld3(C.gc, n, l);
      // Particle::reflect (part.rs:374-383)
      V3<double> loc_r = loc + ((2. * dot(fc - loc, nrm)) * nrm);
      V3<double> cen_r = cen + ((2. * dot(fc - cen, nrm)) * nrm);
      V3<double> gc_r = gc + ((2. * dot(fc - gc, nrm)) * nrm);
      V3<double> dx_face = fc - gc;
      double dx_gradients = len(gc_r - gc);
      V3<double> dx_centroid = cen_r - cen;
      V3<double> dx = loc_r - loc;
      St<double> p = ld5(C.prim, n, l);
      double dtp = C.dt[l];
      double v_max = 0.;
      if (p.c[0] > 0.) v_max += sound_speed(cfg.gd, p.c[4], 1. / p.c[0]);
      Gr<double> gp = ldg(C.grad, n, l);
      St<double> ex = ld5(C.extrap, n, l);
      V3<double> vrel = ld3(C.vrel, n, l);
      St<T> pT = cvtst<T>(p);
      Gr<T> gT;
#pragma unroll
      for (int q = 0; q < 5; ++q) gT.g[q] = cvt3<T>(gp.g[q]);
      V3<T> dxfT = cvt3<T>(dx_face);
      V3<T> nT = cvt3<T>(nrm);
      St<T> dash = ((pT + gdot(gT, dxfT)) + cvtst<T>(ex)) + time_extrapolations(pT, gT, cvt3<T>(vrel), (T)(-0.5 * dtp), g);
      T dx_fac = len(dxfT) / (T)dx_gradients;
      St<T> bstate;
      if (cfg.boundary == BND_REFLECTIVE) {
        St<T> rp = reflect(pT, nT);
        dash = pairwise_limiter(pT, rp, dash, dx_fac, ev);
        St<double> rpd = reflect(p, nrm);
        v_max -= hmin(dot(vel(rpd) - vel(p), dx_centroid), 0.);
        bstate = reflect(dash, nT);
      } else if (cfg.boundary == BND_OPEN) {
        dash = pairwise_limiter(pT, pT, dash, dx_fac, ev);
        bstate = dash;
      } else if (cfg.boundary == BND_VACUUM) {
        St<T> vac = vacuum<T>();
        dash = pairwise_limiter(pT, vac, dash, dx_fac, ev);
        v_max += hmin(dot(vel(p), dx_centroid), 0.);
        bstate = vac;
      } else {
        ev.fail(E_UNSUPPORTED);
        bstate = dash;
      }
      St<T> Fl = riemann_flux(cfg.solver, dash, bstate, mk3(T(0), T(0), T(0)), nT, g, cfg.gd, ev);
      St<double> fluxes = area * cvtst<double>(Fl);
      St<double> rec = dtp * fluxes;
#pragma unroll
      for (int k = 0; k < 5; ++k) outv[k] = rec.c[k];
      outv[5] = v_max;
      outv[6] = area / len(dx);
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) F.rec[k * nf + f] = outv[k];
    flush_events(ev, cnt);
  }
  wave_count(computed, &cnt->n_face_active);
}

// runner apply_fluxes (hydrodynamics.rs:13-39) + kick2 (:51-57) + convert_conserved_to_primitive
// (euler_eqns_fvs.rs:63-69, part.rs:219-261): one thread per cell walks its faces in CSR order, so the
// FP64 accumulation order is the reference's and needs no atomics.
// mode bit 1: accumulate + apply + kick2; mode bit 2: conserved -> primitive.
__global__ void __launch_bounds__(256) k_apply_c2p(Cells C, Faces F, Cfg cfg, DevCounters* cnt, int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n) return;
  const int64_t n = C.n, nf = F.nf;
  Events ev;
  bool a = C.act[i];
  if (!(mode & 1)) {
    if (!a) return;
    St<double> c = ld5(C.cons, n, i);
    St<double> prim;
    if (!(c.c[0] >= 0.)) ev.fail(E_NEG_MASS);
    if (!(c.c[4] >= 0.)) ev.fail(E_NEG_ENERGY);
    if (c.c[0] == 0.) {
      if (c.c[4] != 0.) ev.fail(E_NEG_ENERGY);
      prim = vacuum<double>();
    } else if (c.c[0] > 0.) {
      double m_inv = 1. / c.c[0];
      double density = c.c[0] / C.vol[i];
      V3<double> velocity = m_inv * vel(c);
      double thermal = c.c[4] - 0.5 * len2(vel(c)) * m_inv;
      double pressure = gas_p_from_u(cfg.gd, thermal * m_inv, density);
      if (!(density >= 0.) || !(pressure >= 0.)) ev.fail(E_NEG_PRESSURE);
      prim = mkst(density, velocity, pressure);
    } else {
      prim = vacuum<double>();
    }
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (!hfinite(prim.c[q])) ev.fail(E_NONFINITE);
    st5(C.prim, n, i, prim);
    flush_events(ev, cnt);
    return;
  }
  St<double> acc = ld5(C.facc, n, i);
  double msv = 0., maor = 0.;
  if (a) {
    msv = C.msv[i];
    maor = C.maor[i];
  }
  int32_t k0 = F.off[i], k1 = F.off[i + 1];
  for (int32_t k = k0; k < k1; ++k) {
    int32_t e = F.conn[k];
    int64_t f = e >> 1;
    St<double> rec;
#pragma unroll
    for (int q = 0; q < 5; ++q) rec.c[q] = F.rec[q * nf + f];
    if ((e & 1) == 0)
      acc = acc - rec;
    else
      acc = acc + rec;
    if (a) {
      msv = hmax(F.rec[5 * nf + f], msv);
      maor = hmax(maor, F.rec[6 * nf + f]);
    }
  }
  if (a) {
    St<double> c = ld5(C.cons, n, i) + acc;  // apply_flux (part.rs:178-193)
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (!hfinite(c.c[q])) ev.fail(E_NONFINITE);
    if (c.c[0] < 0.) {
      ev.flux_reset++;
      c = vacuum<double>();
    }
    if (c.c[4] < 0.) {
      ev.flux_reset++;
      c = vacuum<double>();
    }
    acc = vacuum<double>();
    // kick2 -> grav_kick with a_grav = 0 (part.rs:303-312): momentum += (0 * mass)
    c.c[1] += 0.0;
    c.c[2] += 0.0;
    c.c[3] += 0.0;
    st5(C.cons, n, i, c);
    C.msv[i] = msv;
    C.maor[i] = maor;
    if (!(mode & 2)) {
      st5(C.facc, n, i, acc);
      flush_events(ev, cnt);
      return;
    }
    // convert_conserved_to_primitive
    St<double> prim;
    if (!(c.c[0] >= 0.)) ev.fail(E_NEG_MASS);
    if (!(c.c[4] >= 0.)) ev.fail(E_NEG_ENERGY);
    if (c.c[0] == 0.) {
      if (c.c[4] != 0.) ev.fail(E_NEG_ENERGY);
      prim = vacuum<double>();
    } else if (c.c[0] > 0.) {
      double m_inv = 1. / c.c[0];
      double density = c.c[0] / C.vol[i];
      V3<double> velocity = m_inv * vel(c);
      double thermal = c.c[4] - 0.5 * len2(vel(c)) * m_inv;
      double u = thermal * m_inv;
      double pressure = gas_p_from_u(cfg.gd, u, density);
      if (!(density >= 0.) || !(pressure >= 0.)) ev.fail(E_NEG_PRESSURE);
      prim = mkst(density, velocity, pressure);
    } else {
      prim = vacuum<double>();
    }
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (!hfinite(prim.c[q])) ev.fail(E_NONFINITE);
    st5(C.prim, n, i, prim);
  }
  st5(C.facc, n, i, acc);
  flush_events(ev, cnt);
}

// glam::DMat3 as three columns
struct M3 {
  V3<double> x, y, z;
};

// Particle::radius (part.rs:439-447)
__device__ __forceinline__ double cell_radius(double V, const Cfg& cfg) {
  if (cfg.dim == 1) return 0.5 * V;
  if (cfg.dim == 2) return sqrt(0.318309886183790671537767526745028724 * V);
  return pow(cfg.radius3_fac * V, 1. / 3.);
}

// runner gradient_estimate / gradient_limit / gradient_apply (hydrodynamics.rs:67-143, gradients.rs:8-103)
// fused with compute_timesteps (euler_eqns_fvs.rs:50-61, part.rs:42-84, 328-372) and timesteps_apply
// (hydrodynamics.rs:145-158, timeline.rs:103-139) for the active cells.
__global__ void __launch_bounds__(256) k_grad_timestep(Cells C, Faces F, Cfg cfg, DevCounters* cnt, uint64_t ti_current,
                                                       double time_base, double time_base_inv, int do_grad,
                                                       int do_ts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long my_dti = ~0ull;
  if (i < C.n && C.act[i]) {
    const int64_t n = C.n, nf = F.nf;
    Events ev;
    St<double> prim = ld5(C.prim, n, i);
    V3<double> cen = ld3(C.cen, n, i);
    Gr<double> gr;
    V3<double> gcen;
    int32_t k0 = F.off[i], k1 = F.off[i + 1];
    if (do_grad && cfg.runner == 0) {
      // ---- estimate
      Gr<double> g;
#pragma unroll
      for (int q = 0; q < 5; ++q) g.g[q] = mk3(0., 0., 0.);
      M3 m;
      m.x = mk3(0., 0., 0.);
      m.y = mk3(0., 0., 0.);
      m.z = mk3(0., 0., 0.);
      if (cfg.dim == 1) {
        m.y.y = 1.;
        m.z.z = 1.;
      } else if (cfg.dim == 2) {
        m.z.z = 1.;
      }
      for (int32_t k = k0; k < k1; ++k) {
        int32_t e = F.conn[k];
        int64_t f = e >> 1;
        int side = e & 1;
        double area = F.area[f];
        V3<double> sh = decode_shift(F.shc[f], cfg.box);
        V3<double> shift = side == 0 ? sh : -sh;
        int32_t other = side == 0 ? F.right[f] : F.left[f];
        V3<double> oc;
        St<double> op;
        if (other >= 0) {
          oc = ld3(C.cen, n, other);
          op = ld5(C.prim, n, other);
        } else {  // Space::get_boundary_part (space.rs:213-229)
          V3<double> fc = ld3(F.fcen, nf, f), nrm = ld3(F.nrm, nf, f);
          oc = cen + ((2. * dot(fc - cen, nrm)) * nrm);
          op = prim;
          if (cfg.boundary == BND_REFLECTIVE) op = reflect(prim, nrm);
          else if (cfg.boundary == BND_VACUUM) op = vacuum<double>();
        }
        V3<double> ds = (oc + shift) - cen;
        double w = area / len2(ds);
#pragma unroll
        for (int q = 0; q < 5; ++q) g.g[q] = g.g[q] + ((w * (op.c[q] - prim.c[q])) * ds);
        m.x = m.x + (w * (ds.x * ds));
        m.y = m.y + (w * (ds.y * ds));
        m.z = m.z + (w * (ds.z * ds));
      }
      // DMat3::inverse (glam 0.27): cofactors by cross products, det = z . (x cross y), transpose
      V3<double> t0 = cross(m.y, m.z), t1 = cross(m.z, m.x), t2 = cross(m.x, m.y);
      double det = dot(m.z, t2);
      double inv_det = 1. / det;
      V3<double> c0 = mk3(t0.x * inv_det, t0.y * inv_det, t0.z * inv_det);
      V3<double> c1 = mk3(t1.x * inv_det, t1.y * inv_det, t1.z * inv_det);
      V3<double> c2 = mk3(t2.x * inv_det, t2.y * inv_det, t2.z * inv_det);
      M3 inv;
      inv.x = mk3(c0.x, c1.x, c2.x);
      inv.y = mk3(c0.y, c1.y, c2.y);
      inv.z = mk3(c0.z, c1.z, c2.z);
      bool fin = true;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        V3<double> v = g.g[q];
        V3<double> r = v.x * inv.x;
        r = r + (v.y * inv.y);
        r = r + (v.z * inv.z);
        g.g[q] = r;
        fin = fin && finite3(r);
      }
      if (!fin) {
        ev.grad_fallback++;
#pragma unroll
        for (int q = 0; q < 5; ++q) g.g[q] = mk3(0., 0., 0.);
      }
      // ---- limit (LimiterData)
      St<double> lmin = prim, lmax = prim, emin, emax;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        emin.c[q] = INFINITY;
        emax.c[q] = -INFINITY;
      }
      for (int32_t k = k0; k < k1; ++k) {
        int32_t e = F.conn[k];
        int64_t f = e >> 1;
        int side = e & 1;
        V3<double> sh = decode_shift(F.shc[f], cfg.box);
        V3<double> shift = side == 0 ? sh : -sh;
        V3<double> fc = ld3(F.fcen, nf, f);
        St<double> ext = gdot(g, (fc - cen) - shift);
        int32_t other = side == 0 ? F.right[f] : F.left[f];
        St<double> op;
        if (other >= 0) {
          op = ld5(C.prim, n, other);
        } else {
          op = prim;
          if (cfg.boundary == BND_REFLECTIVE) op = reflect(prim, ld3(F.nrm, nf, f));
          else if (cfg.boundary == BND_VACUUM) op = vacuum<double>();
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          lmin.c[q] = hmin(lmin.c[q], op.c[q]);
          lmax.c[q] = hmax(lmax.c[q], op.c[q]);
          emin.c[q] = hmin(emin.c[q], ext.c[q]);
          emax.c[q] = hmax(emax.c[q], ext.c[q]);
        }
      }
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        double d_max = lmax.c[q] - prim.c[q], d_min = lmin.c[q] - prim.c[q];
        double alpha = (emax.c[q] == 0. || emin.c[q] == 0.) ? 1. : hmin(1.0, hmin(d_max / emax.c[q], d_min / emin.c[q]));
        g.g[q] = alpha * g.g[q];
      }
      gr = g;
      gcen = cen;
      stg(C.grad, n, i, gr);
      st3(C.gc, n, i, gcen);
    } else {
      gr = ldg(C.grad, n, i);
      gcen = ld3(C.gc, n, i);
    }
    // ---- Particle::timestep (part.rs:42-84)
    St<double> c = ld5(C.cons, n, i);
    double dt_req = INFINITY;
    if (do_ts) {
    if (c.c[0] == 0.) {
      st3(C.v, n, i, mk3(0., 0., 0.));
      dt_req = INFINITY;
    } else {
      double m_inv = 1. / c.c[0];
      V3<double> loc = ld3(C.loc, n, i);
      V3<double> d_v = vel(gdot(gr, loc - gcen));
      V3<double> fluid_v = (m_inv * vel(c)) + d_v;
      double uint_ = (c.c[4] - 0.5 * dot(vel(c), vel(prim))) / c.c[0];
      double V = C.vol[i];
      double cs = sound_speed(cfg.gd, gas_p_from_u(cfg.gd, uint_, prim.c[0]), V * m_inv);
      V3<double> v;
      switch (cfg.motion) {
        case MOT_FIXED:
          v = mk3(0., 0., 0.);
          break;
        case MOT_FLUID:
          v = fluid_v;
          break;
        case MOT_STEER: {
          V3<double> d = cen - loc;
          double abs_d = len(d);
          double eta_r = 0.25 * cell_radius(V, cfg);
          if (abs_d > 0.9 * eta_r) {
            double fac = 0.99 * cs / abs_d;
            if (abs_d < 1.1 * eta_r) fac *= 5. * (abs_d - 0.9 * eta_r) / eta_r;
            v = fluid_v + (fac * d);
          } else {
            v = fluid_v;
          }
          break;
        }
        default: {  // STEER_PAKMOR
          double alpha = 2. / (double)cfg.dim * C.maor[i];
          double beta = 2.25;
          double eta = hmax(0., hmin(0.5, 2. / beta * (alpha - 0.75 * beta)));
          V3<double> dd = cen - loc;
          double lr = 1. / len(dd);
          V3<double> d = lr * dd;
          V3<double> curl = mk3(gr.g[3].y - gr.g[2].z, gr.g[1].z - gr.g[3].x, gr.g[2].x - gr.g[1].y);
          double v_reg = cell_radius(V, cfg) * len(curl);
          v = fluid_v + ((eta * hmax(cs, v_reg)) * d);
          break;
        }
      }
      if (!finite3(v)) ev.fail(E_NONFINITE);
      V3<double> v_rel = fluid_v - v;
      st3(C.v, n, i, v);
      st3(C.vrel, n, i, v_rel);
      C.maor[i] = 0.;
      double v_max = hmax(C.msv[i], len(v_rel) + cs);
      C.msv[i] = 0.;
      dt_req = v_max > 0. ? cfg.cfl * cell_radius(V, cfg) / v_max : INFINITY;
    }
    // ---- timesteps_apply
    int8_t bin = C.bin[i];
    uint64_t dti = make_integer_timestep(dt_req, bin, HFV_NUM_TIME_BINS, ti_current, time_base_inv);
    C.dt[i] = (double)dti * time_base;
    C.bin[i] = get_time_bin(dti);
    my_dti = dti;
    }
    flush_events(ev, cnt);
  }
  // wave-level min, then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(my_dti, off);
    my_dti = o < my_dti ? o : my_dti;
  }
  if (__lane_id() == 0 && my_dti != ~0ull) atomicMin(&cnt->ti_min, my_dti);
}

// timestep_limiter, first sweep (hydrodynamics.rs:165-179): wakeup from active neighbours' new bins.
__global__ void __launch_bounds__(256) k_wakeup(Cells C, Faces F, DevCounters* cnt, int sync_all) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int mybin = 127;
  if (i < C.n) {
    int8_t w = C.bin[i];
    mybin = w;
    if (!sync_all) {
      int32_t k0 = F.off[i], k1 = F.off[i + 1];
      for (int32_t k = k0; k < k1; ++k) {
        int32_t e = F.conn[k];
        int64_t f = e >> 1;
        int32_t other = (e & 1) == 0 ? F.right[f] : F.left[f];
        if (other < 0) continue;
        if (!C.act[other]) continue;
        int8_t cand = (int8_t)(C.bin[other] + 2);
        if (cand < w) w = cand;
      }
      C.wake[i] = w;
    }
  }
  if (sync_all) {  // timestep_sync (hydrodynamics.rs:186-191): min bin over all cells
    for (int off = 32; off > 0; off >>= 1) {
      int o = __shfl_xor(mybin, off);
      mybin = o < mybin ? o : mybin;
    }
    if (__lane_id() == 0 && mybin != 127) atomicMin(&cnt->min_bin, mybin);
  }
}

// Particle::timestep_limit (part.rs:263-297) + reset_extrapolations (hydrodynamics.rs:193-195) +
// kick1 (hydro_kick1 part.rs:87-104, grav_kick with a_grav = 0).
__global__ void __launch_bounds__(256) k_limit_kick(Cells C, Cfg cfg, DevCounters* cnt, int sync_all,
                                                    uint64_t ti_current, double time_base) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n) return;
  const int64_t n = C.n;
  int8_t wakeup = sync_all ? (int8_t)cnt->min_bin : C.wake[i];
  int8_t bin = C.bin[i];
  if (wakeup < bin) {
    double dt = C.dt[i];
    if (bin <= get_max_active_bin(ti_current)) {
      bin = wakeup;
      dt = (double)get_integer_timestep(bin) * time_base;
    } else {
      uint64_t ti_end_old = get_integer_time_end(ti_current, bin);
      dt -= (double)(ti_end_old - ti_current) * time_base;
      bin = wakeup;
      uint64_t dti_new = get_integer_timestep(bin);
      uint64_t ti_end = get_integer_time_end(ti_current, bin);
      dt += (ti_end == ti_current) ? (double)dti_new * time_base : (double)(ti_end - ti_current) * time_base;
    }
    C.bin[i] = bin;
    C.dt[i] = dt;
  }
  if (C.act[i]) {
    st5(C.extrap, n, i, vacuum<double>());
    if (cfg.motion != MOT_FIXED) {
      double rho = C.prim[i];
      if (rho > 0.) {
        double kick_fac = 0.5 * C.dt[i] / rho;
        for (int a = 0; a < cfg.dim; ++a) C.v[a * n + i] -= kick_fac * C.grad[(12 + a) * n + i];
      }
    }
    C.cons[1 * n + i] += 0.0;
    C.cons[2 * n + i] += 0.0;
    C.cons[3 * n + i] += 0.0;
  }
}

// Space::drift (space.rs:833-854) with Particle::drift / box_wrap / box_reflect (part.rs:107-119, 449-469).
__global__ void k_drift(Cells C, Cfg cfg, double dt, DevCounters* cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n) return;
  const int64_t n = C.n;
  const int dim = cfg.dim;
  Events ev;
  double loc[3], cen[3], gc[3];
  for (int a = 0; a < 3; ++a) {
    loc[a] = C.loc[a * n + i];
    cen[a] = C.cen[a * n + i];
    gc[a] = C.gc[a * n + i];
  }
  for (int a = 0; a < dim; ++a) {
    double va = C.v[a * n + i];
    loc[a] += va * dt;
    cen[a] += va * dt;
    gc[a] += va * dt;
  }
  if (!(isfinite(loc[0]) && isfinite(loc[1]) && isfinite(loc[2]))) ev.fail(E_DRIFT);
  bool inside = true;
  for (int a = 0; a < dim; ++a) inside = inside && (loc[a] >= 0. && loc[a] < cfg.box[a]);
  if (!inside) {
    if (cfg.boundary == BND_PERIODIC) {
      double old[3] = {loc[0], loc[1], loc[2]};
      for (int a = 0; a < dim; ++a) {
        int guard = 0;
        while (loc[a] < 0. && ++guard < 64) loc[a] += cfg.box[a];
        while (loc[a] >= cfg.box[a] && ++guard < 128) loc[a] -= cfg.box[a];
      }
      for (int a = 0; a < 3; ++a) {
        double s = loc[a] - old[a];
        gc[a] += s;
        cen[a] += s;
      }
    } else if (cfg.boundary == BND_REFLECTIVE) {
      double old[3] = {loc[0], loc[1], loc[2]};
      for (int a = 0; a < dim; ++a) {
        if (loc[a] < 0.) loc[a] -= 2. * loc[a];
        if (loc[a] > cfg.box[a]) loc[a] -= 2. * (loc[a] - cfg.box[a]);
      }
      bool in2 = true;
      for (int a = 0; a < dim; ++a) in2 = in2 && (loc[a] >= 0. && loc[a] < cfg.box[a]);
      if (!in2) ev.fail(E_DRIFT);
      V3<double> sh = mk3(loc[0] - old[0], loc[1] - old[1], loc[2] - old[2]);
      if (len2(sh) > 0.) {
        V3<double> nn = (1. / len(sh)) * sh;
        St<double> p = ld5(C.prim, n, i);
        st5(C.prim, n, i, reflect(p, nn));
        // Particle::reflect_gradients (part.rs:390-429), including its use of the already reflected
        // density gradient for the pressure gradient
        Gr<double> g = ldg(C.grad, n, i);
        g.g[0] = g.g[0] - ((2. * dot(g.g[0], nn)) * nn);
        g.g[4] = g.g[4] - ((2. * dot(g.g[0], nn)) * nn);
        V3<double> dv_dx = mk3(g.g[1].x, g.g[2].x, g.g[3].x);
        V3<double> dv_dy = mk3(g.g[1].y, g.g[2].y, g.g[3].y);
        V3<double> dv_dz = mk3(g.g[1].z, g.g[2].z, g.g[3].z);
        V3<double> dvn = mk3(dot(dv_dx, nn), dot(dv_dy, nn), dot(dv_dz, nn));
        g.g[1] = g.g[1] - (nn.x * (2. * dvn));
        g.g[2] = g.g[2] - (nn.y * (2. * dvn));
        g.g[3] = g.g[3] - (nn.z * (2. * dvn));
        for (int q = 1; q <= 3; ++q) g.g[q] = g.g[q] - ((2. * dot(g.g[q], nn)) * nn);
        stg(C.grad, n, i, g);
      }
    } else {
      ev.fail(E_UNSUPPORTED);  // particle removal at open/vacuum walls is not on this path
    }
  }
  for (int a = 0; a < 3; ++a) {
    C.loc[a * n + i] = loc[a];
    C.cen[a * n + i] = cen[a];
    C.gc[a * n + i] = gc[a];
  }
  flush_events(ev, cnt);
}

// 1D Voronoi rebuild on the device for moving meshes (Space::volume_calculation, space.rs:236-263, 1D
// semantics pinned by test_init_1d, space.rs:1453-1491): generators stay sorted, so the face topology is
// fixed and only face positions, volumes and centroids change.  Interior face k sits between cells
// k-1 and k; non-periodic meshes have walls at faces 0 and N; periodic meshes wrap at face 0.
__global__ void k_voronoi1d_faces(Cells C, Faces F, Cfg cfg, int periodic, DevCounters* cnt) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= F.nf) return;
  const int64_t n = C.n;
  double x;
  if (k >= 1 && k < n) {
    double a = C.loc[k - 1], b = C.loc[k];
    if (!(a < b)) atomicCAS(&cnt->err, 0, E_UNSUPPORTED);
    x = 0.5 * (a + b);
  } else if (!periodic) {
    x = (k == 0) ? 0. : cfg.box[0];
  } else {
    x = 0.5 * (C.loc[n - 1] + (C.loc[0] + cfg.box[0]));
  }
  F.fcen[k] = x;
}
__global__ void k_voronoi1d_cells(Cells C, Faces F, Cfg cfg, int periodic) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n || !C.act[i]) return;
  const int64_t n = C.n;
  double xl, xr;
  if (!periodic) {
    xl = F.fcen[i];
    xr = F.fcen[i + 1];
  } else {
    xl = (i == 0) ? F.fcen[0] - cfg.box[0] : F.fcen[i];
    xr = (i == n - 1) ? F.fcen[0] : F.fcen[i + 1];
  }
  double V = xr - xl;
  C.vol[i] = V;
  C.cen[i] = 0.5 * (xl + xr);
  C.cen[n + i] = 0.;
  C.cen[2 * n + i] = 0.;
  C.srad[i] = 1.5 * V;
}

// Particle::from_ic + first_init (part.rs:200-217)
__global__ void k_init_ic(Cells C, Cfg cfg, const double* mass, const double* vel3, const double* u_int,
                          DevCounters* cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n) return;
  const int64_t n = C.n;
  Events ev;
  double m = mass[i];
  V3<double> v = ld3(vel3, n, i);
  St<double> c = mkst(m, m * v, m * u_int[i] + 0.5 * m * len2(v));
  St<double> prim;
  if (c.c[0] > 0.) {
    double m_inv = 1. / c.c[0];
    double density = c.c[0] / C.vol[i];
    V3<double> velocity = m_inv * vel(c);
    double thermal = c.c[4] - 0.5 * len2(vel(c)) * m_inv;
    double pressure = gas_p_from_u(cfg.gd, thermal * m_inv, density);
    if (!(density >= 0.) || !(pressure >= 0.)) ev.fail(E_NEG_PRESSURE);
    prim = mkst(density, velocity, pressure);
  } else {
    prim = vacuum<double>();
  }
  st5(C.cons, n, i, c);
  st5(C.prim, n, i, prim);
  flush_events(ev, cnt);
}

// Space::status sums (space.rs:1248-1280): one block, fixed-order tree reduction per component.
__global__ void __launch_bounds__(1024) k_conservation(Cells C, double* out5) {
  __shared__ double sh[1024];
  const int64_t n = C.n;
  for (int q = 0; q < 5; ++q) {
    double s = 0.;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += C.cons[q * n + i] + C.facc[q * n + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out5[q] = sh[0];
    __syncthreads();
  }
}

// RiemannFluxSolver::solve_for_flux over a batch (unit parity entry point).
template <class T>
__global__ void k_riemann_batch(Cfg cfg, int64_t nb, const double* WL, const double* WR, const double* vf,
                                const double* nr, double* Fo, DevCounters* cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  Events ev;
  St<T> l = cvtst<T>(ld5(WL, nb, i)), r = cvtst<T>(ld5(WR, nb, i));
  V3<T> iv = cvt3<T>(ld3(vf, nb, i)), n = cvt3<T>(ld3(nr, nb, i));
  St<T> f = riemann_flux(cfg.solver, l, r, iv, n, gas_for<T>(cfg), cfg.gd, ev);
  st5(Fo, nb, i, cvtst<double>(f));
  flush_events(ev, cnt);
}

}  // namespace hfv
