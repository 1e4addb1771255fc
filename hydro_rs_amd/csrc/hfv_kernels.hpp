// Device kernels of the per-face finite-volume step (gfx950).  Included by hfv_api.hip.
//
// One launch sequence advances every simulation of a batch by one of its own engine steps (a batch
// of one is the single-simulation path).  Engine state lives in HBM, so steps are enqueued (or replayed
// from a captured graph) without a host round trip.
//
// HBM layout (cells/faces of all simulations concatenated, indices global):
//   cells, component-major SoA (field[c * N + i]): prim[5], grad[15], gc[3], cons[5], facc[5] (pending
//     flux accumulator), vol, loc[3], cen[3], v[3], vrel[3], maor, msv, dt, srad, bin (int8), wake
//     (int8), act (uint8), tex[5] (time-extrapolation rate of prim, grad and v_rel, written by k_grad
//     when they change); extrap[N][5] (AoS: read and written whole by predict, staged whole by the flux
//     kernel)
//     pc[N][8] (AoS: prim[5], cen[3] of every cell after this step's c2p, one 64-byte line per cell
//     for the gradient sweeps' neighbour gathers)
//     fr[N][32] (AoS flux record of static-geometry cells, HFV_FR_*: prim, time-extrapolation rate,
//     sound speed, gradients, dt, gc; two 128-byte lines staged whole by the flux kernel)
//   faces, SoA: area, nrm[3], fcen[3], left (int32), right (int32, -1 = wall), shc (uint8 periodic
//     shift code, 2 bits per axis), cur (uint8: record of this step); rec[F][8] (AoS, one 64-byte line
//     per face: dt*F[5], v_max, a_over_r, unused)
//   csr: off[N+1] (int32), conn[2F] (int32: face << 1 | side, side 1 = the cell is the face's right),
//     nbr[2F] (int32: the face's other cell, -1 for a wall)
//   staging lists: per face chunk its distinct cells (scell, scnt), per face its cells' list slots
//     (floc)
//   block tables: one {sim, lo, hi, chunk} per thread block of the cell kernels and of the face
//     kernels.  A block only holds items of one simulation, so configuration and engine loads are
//     uniform (scalar) and a finished simulation's blocks exit at once.  Blocks are ordered so that
//     the 8 XCDs (round-robin dispatch) each receive a contiguous range of chunks (L2 locality of the
//     neighbour gathers).  Per-chunk partial counts replace global atomics; k_end reduces them.
//
// Step kernels: k_begin (per simulation) -> [k_drift] -> k_mask (cells: active flag, predict) ->
// [k_mesh1d_*] -> k_flux (faces: gate + flux) -> k_apply (cells: accumulate in CSR order, apply, c2p)
// -> k_grad (active cells: gradients, limiter, timestep) -> k_wakeup, k_limit (cells) -> k_end (per
// simulation: partial sums, next sync point).
#pragma once
#include "hfv_physics.hpp"

namespace hfv {

struct Cells {
  int64_t n;
  double *prim, *grad, *gc, *extrap, *cons, *facc, *vol, *loc, *cen, *v, *vrel, *maor, *msv, *dt, *srad;
  double* pc;
  double* tex;   // [5][N] time-extrapolation rate of the cell's prim/grad/v_rel (written when they change)
  uint64_t* tiex;  // [N] FP32-flux contexts: the predict clock (SimEngine::ti_pred) at which extrap is due
  // extrap, tex and grad hold floats (same layouts) in FP32-flux contexts: ld_ex/st_ex/ld_tex/st_tex/
  // ld_grad/st_grad
  double* minv;  // [9][N] inverse of the gradient matrix (static geometry only)
  double* fr;    // [N][HFV_FR_W] flux inputs of static-geometry cells (layout: HFV_FR_*)
  double* dtq;   // [N] requested timesteps of the trait-level compute_timesteps (before timesteps_apply)
  int8_t *bin, *wake;
  uint8_t* act;
};

struct Faces {
  int64_t nf;
  double *area, *nrm, *fcen, *rec;
  int32_t *left, *right, *off, *conn;
  int32_t* nbr;  // per CSR entry: the face's other cell (-1: wall)
  uint8_t *shc, *cur;
  double* eg;    // [2F][8] per CSR entry: w, ds[3], dxf[3] (static gradient geometry only)
  double* fs;    // [F][4] per face: dx_centroid[3], a_over_r (static geometry only)
  // LDS staging of the flux kernel: per face chunk, its distinct cells (scell[chunk][HFV_STAGE_MAX],
  // ascending) and their count (scnt, 0 = more than HFV_STAGE_MAX: no staging); per face, the local
  // indices of its left and right cell in that list (floc[2F], 255 for a wall).
  int32_t* scell;
  uint8_t *scnt, *floc;
  // LDS staging of the gradient sweeps: per cell chunk the distinct neighbour cells of its cells
  // (gcell[chunk][HFV_GSTAGE_MAX], ascending; gcnt = 0: more than HFV_GSTAGE_MAX), and per CSR entry the
  // neighbour's list slot (nloc[2F], 0xFFFF for a wall)
  int32_t* gcell;
  uint16_t *gcnt, *nloc;
  // per cell chunk the other cell chunks its cells' neighbours lie in (nch[chunk][HFV_NCH_MAX]) and their
  // count (nchn; 0xFF: more than HFV_NCH_MAX, or a mesh that changes on the device)
  int32_t* nch;
  uint8_t* nchn;
};

struct Cfg {
  int dim, boundary, motion, runner, sync_all, precision;
  double box[3];
  double cfl;
  double radius3_fac;  // 0.25 * 3. * FRAC_1_PI
  SolverCfg solver;
  Gas<double> gd;
  Gas<float> gf;
  int gravity;  // HFV_GRAV_*
  double grav_soft;
  double grav_pos[3];
  double grav_acc[3];
};

struct SimDesc {
  Cfg cfg;
  int64_t cell_base, n, face_base, nf;
  int32_t cchunk0, cchunk1, fchunk0, fchunk1;  // chunk ranges of the cell / face block tables
  int moving1d;  // 1D moving mesh rebuilt on the device
  int periodic;
  int static_geom;  // fixed mesh with zero particle velocities: face geometry precomputed (k_geom_faces)
  int static_grad;  // ... and gradient geometry too (k_geom_cells; chosen for 3D meshes)
  int staged;       // static_geom and some face chunk has a staging list: the flux records fr are kept
  int gstaged;      // static_geom and the gradient sweeps stage their chunks' neighbour lines
  // 2D/3D moving mesh rebuilt on the device (hfv_voronoi2d.hpp, hfv_voronoi3d.hpp; vmesh = the dimension,
  // 0: none): the bin grid of the tessellator (nb bins of width w per axis; 1 bin of width 1 along an
  // axis past the dimension), the slot's CSR base and capacities, and its count of CSR offsets
  int vmesh;
  int vnb[3];
  int noff;
  int pad_v;
  double vw[3];
  int64_t conn_base, conn_cap, face_cap;
};

// Thread block -> work: simulation (-1: nothing), items [lo, hi), chunk (index of its partials).
struct BlockJob {
  int32_t sim, lo, hi, chunk;
};

#define HFV_MAX_SUBDRIFTS 8
#define HFV_NCH_MAX 8
// Engine of one simulation (src/engine.rs:33-101, 152-268).
struct SimEngine {
  uint64_t ti_old, ti_current, ti_next;
  uint64_t ti_snap, between, dump_lo, dump_hi;  // pending dump sync points: dump_lo, +between, .. < dump_hi
  double time_base, time_base_inv, t_current, t_end;
  double dt_pred;
  double sub_dt[HFV_MAX_SUBDRIFTS];
  int n_sub;
  int stepping;  // this step runs for this simulation
  int done;
  int force;     // next k_begin runs a step at ti_current without the engine's time loop (hfv_step)
  int8_t mab;
  int8_t pad8[3];
  int min_bin;
  unsigned long long ti_min, n_active, n_face_active;
  long long step_count, cell_updates, face_fluxes;
  long long last_active, last_face_active;
  int err;
  // FP32-flux contexts defer the predict: a cell's extrapolations are extrap + (ti_pred - tiex) *
  // time_base * (-rate), with the rate constant between the tag and now.  ti_pred advances by every
  // engine step's predict interval (k_begin), not by the trait-level phases (no_pred), whose predict
  // settles the sums itself.  chk: host-written primitives wait for predict's check_physical.
  int chk;
  int no_pred;
  int pred_step;
  uint64_t ti_pred;
  int n_down;  // 1D periodic moving meshes: faces of this step's rebuild where the generator order steps down
  int nf_mesh;  // 2D moving meshes: faces of the current mesh (the device rebuild's count)
  int nf_tmp;   // 2D moving meshes: the rebuild's provisional face cursor
  int pad_;
};

// Batch-global counters (soft events; the first hard error of any simulation).
struct DevCounters {
  int err;
  unsigned int nfall;  // faces queued for the exact-solver fallback this step
  unsigned long long check_physical, flux_reset, grad_fallback, brent;
};

// Per-chunk partials written by the dense kernels (one entry per block-table chunk).
struct Partials {
  uint32_t* act;    // active cells of a cell chunk (k_mask)
  uint32_t* face;   // computed faces of a face chunk (k_flux)
  uint64_t* ti;     // minimum new dti of a cell chunk's active cells (k_grad)
  uint32_t* cflag;  // a cell chunk has work in the accumulation and limiter sweeps this step (k_apply)
};

struct Batch {
  Cells C;
  Faces F;
  const SimDesc* sims;
  SimEngine* eng;
  int nsims;
  const BlockJob* cjobs;  // cell kernels: grid = number of entries
  const BlockJob* fjobs;  // face kernels
  Partials P;
  DevCounters* cnt;
  int2* fall;  // (face, simulation) queued for k_flux_fallback
};

// Flux record fr[i] of a staged static-geometry cell: the flux phase's per-cell inputs apart from the
// extrapolations, in the context's flux precision.  Written where the fields change (k_grad for
// active cells, k_limit for woken cells' dt, k_mask for check_physical resets, k_pack_fr after host
// changes).  FP64 flux: 32 doubles (two 128-byte lines).
#define HFV_EX_W 5  // extrapolations: AoS extrap[N][5]
#define HFV_FR_W 32     // doubles allocated per cell
#define HFV_FR_DT 0
#define HFV_FR_GC 1     // gc[3]
#define HFV_FR_PRIM 4   // prim[5]
#define HFV_FR_TEX 9    // tex[5]: time-extrapolation rate of prim
#define HFV_FR_CS 14    // sound speed of prim (0 for vacuum)
#define HFV_FR_GRAD 15  // grad[15]
#define HFV_FR_USED 30
// FP32 flux: one 128-byte line of 32-bit words: dt in FP64 (words 0-1), the other fields rounded to
// FP32 (the values the FP32 path computes with; the gathered path rounds the same fields).
#define HFV_FR32_GC 2
#define HFV_FR32_PRIM 5
#define HFV_FR32_TEX 10
#define HFV_FR32_CS 15
#define HFV_FR32_GRAD 16
#define HFV_FR32_USED 31
// LDS staging records: the used fields, then the extrapolations[5] (doubles / words); odd strides, so
// that neighbouring records start in different banks
#define HFV_FR_EX HFV_FR_USED
#define HFV_FR_LDS 35
#define HFV_FR32_EX HFV_FR32_USED
#define HFV_FR32_EL (HFV_FR32_EX + 5)  // the cell's unsummed predict time (float)
#define HFV_FR32_LDS 37
static_assert(HFV_FR32_EL < HFV_FR32_LDS, "staged FP32 record words");
#define HFV_STAGE_MAX 128
#define HFV_GSTAGE_MAX 600  // neighbour pc lines a gradient block stages (64 bytes each)

#define HFV_NUM_TIME_BINS 56
#define HFV_MAX_NR_TIMESTEPS (1ull << 57)
#define HFV_BLOCK 256

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
// Cell and face indices are below 2^29 (checked at allocation), so gathers use a 32-bit byte offset
// from a uniform base: one VGPR per address instead of a 64-bit pair per field.
__device__ __forceinline__ double ldo(const double* a, uint32_t i) {
  return *(const double*)((const char*)a + (uint64_t)(i * 8u));
}
__device__ __forceinline__ V3<double> ld3(const double* a, int64_t n, int64_t i) {
  uint32_t u = (uint32_t)i;
  return mk3(ldo(a, u), ldo(a + n, u), ldo(a + 2 * n, u));
}
__device__ __forceinline__ void st3(double* a, int64_t n, int64_t i, V3<double> v) {
  a[i] = v.x;
  a[n + i] = v.y;
  a[2 * n + i] = v.z;
}
__device__ __forceinline__ St<double> ld5(const double* a, int64_t n, int64_t i) {
  St<double> s;
  uint32_t u = (uint32_t)i;
#pragma unroll
  for (int k = 0; k < 5; ++k) s.c[k] = ldo(a + k * n, u);
  return s;
}
__device__ __forceinline__ void st5(double* a, int64_t n, int64_t i, const St<double>& s) {
#pragma unroll
  for (int k = 0; k < 5; ++k) a[k * n + i] = s.c[k];
}
__device__ __forceinline__ Gr<double> ldg(const double* a, int64_t n, int64_t i) {
  Gr<double> g;
  uint32_t u = (uint32_t)i;
#pragma unroll
  for (int k = 0; k < 5; ++k)
    g.g[k] = mk3(ldo(a + (3 * k) * n, u), ldo(a + (3 * k + 1) * n, u), ldo(a + (3 * k + 2) * n, u));
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
// One 64-byte line of the pc pack: prim[5], cen[3].
__device__ __forceinline__ void store_pc(double* pc, int64_t i, const St<double>& p, const V3<double>& c) {
  double2* d = (double2*)(pc + 8 * (uint64_t)i);
  d[0] = make_double2(p.c[0], p.c[1]);
  d[1] = make_double2(p.c[2], p.c[3]);
  d[2] = make_double2(p.c[4], c.x);
  d[3] = make_double2(c.y, c.z);
}
__device__ __forceinline__ void load_pc(const double* pc, uint32_t i, St<double>& p, V3<double>& c) {
  const double2* d = (const double2*)(pc + 8 * (uint64_t)i);
  double2 a = d[0], b = d[1], e = d[2], g = d[3];
  p.c[0] = a.x;
  p.c[1] = a.y;
  p.c[2] = b.x;
  p.c[3] = b.y;
  p.c[4] = e.x;
  c = mk3(e.y, g.x, g.y);
}

// The flux record of a cell as HFV_FR_W values.
__device__ __forceinline__ void fill_fr(double* v, const St<double>& p, const St<double>& rate, double cs,
                                        const Gr<double>& g, double dt, const V3<double>& gc) {
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    v[HFV_FR_PRIM + q] = p.c[q];
    v[HFV_FR_TEX + q] = rate.c[q];
    v[HFV_FR_GRAD + 3 * q] = g.g[q].x;
    v[HFV_FR_GRAD + 3 * q + 1] = g.g[q].y;
    v[HFV_FR_GRAD + 3 * q + 2] = g.g[q].z;
  }
  v[HFV_FR_CS] = cs;
  v[HFV_FR_DT] = dt;
  v[HFV_FR_GC] = gc.x;
  v[HFV_FR_GC + 1] = gc.y;
  v[HFV_FR_GC + 2] = gc.z;
  v[HFV_FR_USED] = 0.;
  v[HFV_FR_USED + 1] = 0.;
}
// The FP32 flux record as its 16 doubles' worth of bits.
__device__ __forceinline__ void fill_fr32(double* v, const St<double>& p, const St<double>& rate, double cs,
                                          const Gr<double>& g, double dt, const V3<double>& gc) {
  float w[32];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    w[HFV_FR32_PRIM + q] = (float)p.c[q];
    w[HFV_FR32_TEX + q] = (float)rate.c[q];
    w[HFV_FR32_GRAD + 3 * q] = (float)g.g[q].x;
    w[HFV_FR32_GRAD + 3 * q + 1] = (float)g.g[q].y;
    w[HFV_FR32_GRAD + 3 * q + 2] = (float)g.g[q].z;
  }
  w[HFV_FR32_CS] = (float)cs;
  w[HFV_FR32_GC] = (float)gc.x;
  w[HFV_FR32_GC + 1] = (float)gc.y;
  w[HFV_FR32_GC + 2] = (float)gc.z;
  w[HFV_FR32_USED] = 0.f;
  v[0] = dt;
#pragma unroll
  for (int k = 1; k < 16; ++k) v[k] = __hiloint2double(__float_as_int(w[2 * k + 1]), __float_as_int(w[2 * k]));
#pragma unroll
  for (int k = 16; k < HFV_FR_W; ++k) v[k] = 0.;
}
// The record in the context's precision: doubles of the record (16 of them carry the FP32 record).
__device__ __forceinline__ int fill_fr_any(double* v, bool f32, const St<double>& p, const St<double>& rate, double cs,
                                           const Gr<double>& g, double dt, const V3<double>& gc) {
  if (f32) {
    fill_fr32(v, p, rate, cs, g, dt, gc);
    return 16;
  }
  fill_fr(v, p, rate, cs, g, dt, gc);
  return HFV_FR_W;
}
// The flux record of cell i (one or two 128-byte lines).
__device__ __forceinline__ void store_fr(double* fr, int64_t i, bool f32, const St<double>& p, const St<double>& rate,
                                         double cs, const Gr<double>& g, double dt, const V3<double>& gc) {
  double v[HFV_FR_W];
  const int nd = fill_fr_any(v, f32, p, rate, cs, g, dt, gc);
  double2* d = (double2*)(fr + HFV_FR_W * (uint64_t)i);
#pragma unroll
  for (int q = 0; q < HFV_FR_W / 2; ++q)
    if (2 * q < nd) d[q] = make_double2(v[2 * q], v[2 * q + 1]);
}
// A wave's new flux records (lane l: cell i0 + l, when bit l of `has` is set), transposed through the
// wave's own LDS buffer buf[64 * 17] in two halves, so that every store instruction writes whole
// 128-byte lines.  Wave-level only (no block barrier): LDS operations of a wave complete in order.
__device__ __forceinline__ void store_fr_wave(double* fr, int64_t i0, uint64_t has, const double* v, double* buf,
                                              int halves) {
  const int lane = (int)__lane_id();
  const bool mine = (has >> lane) & 1ull;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h >= halves) break;
    if (mine) {
#pragma unroll
      for (int j = 0; j < 16; ++j) buf[lane * 17 + j] = v[16 * h + j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int k = lane + 64 * it, rr = k >> 3, q = k & 7;
      if ((has >> rr) & 1ull)
        ((double2*)(fr + HFV_FR_W * (uint64_t)(i0 + rr) + 16 * h))[q] =
            make_double2(buf[rr * 17 + 2 * q], buf[rr * 17 + 2 * q + 1]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}
// Extrapolations (AoS [N][5]) and kept rates (SoA [5][N]) in the context's flux precision: doubles in
// FP64 contexts, floats in the same layouts in FP32-flux contexts (reconstruction inputs; the FP32 path
// rounds them to FP32 where it uses them anyway, so storing them rounded changes nothing downstream
// except the FP32 rounding of the stored extrapolation sum at each step).
__device__ __forceinline__ St<double> ld_ex(const Cells& C, int64_t i, bool f32) {
  if (f32) {
    const float* e = (const float*)C.extrap + HFV_EX_W * (uint64_t)i;
    return mkst((double)e[0], mk3((double)e[1], (double)e[2], (double)e[3]), (double)e[4]);
  }
  const double* e = C.extrap + HFV_EX_W * (uint64_t)i;
  return mkst(e[0], mk3(e[1], e[2], e[3]), e[4]);
}
__device__ __forceinline__ void st_ex(const Cells& C, int64_t i, bool f32, const St<double>& v) {
  if (f32) {
    float* e = (float*)C.extrap + HFV_EX_W * (uint64_t)i;
#pragma unroll
    for (int q = 0; q < 5; ++q) e[q] = (float)v.c[q];
    return;
  }
  double* e = C.extrap + HFV_EX_W * (uint64_t)i;
#pragma unroll
  for (int q = 0; q < 5; ++q) e[q] = v.c[q];
}
__device__ __forceinline__ St<double> ld_tex(const Cells& C, int64_t i, bool f32) {
  if (f32) {
    const float* t = (const float*)C.tex;
    const uint64_t n = (uint64_t)C.n;
    return mkst((double)t[i], mk3((double)t[n + i], (double)t[2 * n + i], (double)t[3 * n + i]), (double)t[4 * n + i]);
  }
  return ld5(C.tex, C.n, i);
}
__device__ __forceinline__ void st_tex(const Cells& C, int64_t i, bool f32, const St<double>& v) {
  if (f32) {
    float* t = (float*)C.tex;
#pragma unroll
    for (int q = 0; q < 5; ++q) t[(uint64_t)q * (uint64_t)C.n + i] = (float)v.c[q];
    return;
  }
  st5(C.tex, C.n, i, v);
}

// Gradients (SoA [15][N]) in the context's flux precision; FP32-flux contexts keep them rounded to FP32
// (k_grad rounds its result before any use, so every consumer sees the stored values).
__device__ __forceinline__ Gr<double> ld_grad(const Cells& C, int64_t i, bool f32) {
  if (f32) {
    const float* a = (const float*)C.grad;
    const uint64_t n = (uint64_t)C.n;
    Gr<double> g;
#pragma unroll
    for (int k = 0; k < 5; ++k)
      g.g[k] = mk3((double)a[(3 * k) * n + i], (double)a[(3 * k + 1) * n + i], (double)a[(3 * k + 2) * n + i]);
    return g;
  }
  return ldg(C.grad, C.n, i);
}
__device__ __forceinline__ void st_grad(const Cells& C, int64_t i, bool f32, const Gr<double>& g) {
  if (f32) {
    float* a = (float*)C.grad;
    const uint64_t n = (uint64_t)C.n;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      a[(3 * k) * n + i] = (float)g.g[k].x;
      a[(3 * k + 1) * n + i] = (float)g.g[k].y;
      a[(3 * k + 2) * n + i] = (float)g.g[k].z;
    }
    return;
  }
  stg(C.grad, C.n, i, g);
}
__device__ __forceinline__ Gr<double> round_grad(const Gr<double>& g, bool f32) {
  if (!f32) return g;
  Gr<double> r;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.g[k] = mk3((double)(float)g.g[k].x, (double)(float)g.g[k].y, (double)(float)g.g[k].z);
  return r;
}

// The flux phase's per-cell rate and sound speed of primitives p (k_mask's expressions).
__device__ __forceinline__ St<double> fr_rate(const St<double>& p, const Gr<double>& g, const V3<double>& vrel,
                                              const Gas<double>& gd) {
  return p.c[0] > 0. ? tex_rate(p, g, vrel, gd) : vacuum<double>();
}
__device__ __forceinline__ double fr_cs(const St<double>& p, const Gas<double>& gd) {
  return p.c[0] > 0. ? sound_speed(gd, p.c[4], 1. / p.c[0]) : 0.;
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

// Keeps the compiler from moving memory operations across this point: the flux and gradient kernels
// issue their gathers stage by stage, so that a stage's loads are not all in flight (and holding
// registers) at once.
#define HFV_STAGE() asm volatile("" ::: "memory")
// The flux kernel's stage barriers can be compiled out (HFV_FLUX_STAGES 0) to let the scheduler overlap
// the stages' gathers at the cost of registers.  Static-geometry FP32 HLLC faces run without them (the
// 1D/2D ensemble's flux -1.1%; the 3D AIRS flux was 1-2% slower without them, FP64 and FP32)
#ifndef HFV_FLUX_STAGES
#define HFV_FLUX_STAGES 1
#endif
template <class T, int SOLVER, bool STATIC>
constexpr bool flux_stages() {
  return HFV_FLUX_STAGES && !(STATIC && sizeof(T) == 4 && SOLVER == SOLVER_HLLC);
}
#define HFV_FSTAGE()                                  \
  do {                                                \
    if constexpr (flux_stages<T, SOLVER, STATIC>()) { \
      HFV_STAGE();                                    \
    }                                                 \
  } while (0)
// Minimum resident blocks per CU (of 4 waves, one per SIMD) the compiler must fit the flux / gradient
// kernels into (0: its choice).  Static-geometry FP32 flux: 6, where the one-chunk variants fit 6 waves
// per SIMD (76-78 VGPRs, no spills): 3D FP32 flux -10% against 5 waves (96 VGPRs), which was itself 12%
// faster than the compiler's choice; 7 for its AIRS variants (72 VGPRs, 12 spilled on cold paths): 3D
// FP32 AIRS flux -6% more, where the 1D HLLC flux of the ensemble lost 0.4% at 7.  The others: 5, which
// the per-step-geometry variants meet (93-96 VGPRs; at 6 they spill more); the staged FP64 ones keep 4
// (their LDS allows no more).
#ifndef HFV_FLUX_WAVES_F32
#define HFV_FLUX_WAVES_F32 6
#endif
#ifndef HFV_FLUX_WAVES_F32_AIRS
#define HFV_FLUX_WAVES_F32_AIRS 7
#endif
#ifndef HFV_FLUX_WAVES
#define HFV_FLUX_WAVES 5
#endif
// CSR entries of a cell whose loads the per-step-geometry gradient sweeps issue together: 4 against a
// two-entry software pipeline, KH 128^3 gradient phase -8% (FP64) / -9% (FP32 flux), ensemble -5%; 8
// costs a wave per SIMD (183 VGPRs) and was slower
#ifndef HFV_GRAD_CHUNK
#define HFV_GRAD_CHUNK 4
#endif
// CSR entries of a cell whose loads (CSR value, face flag, record) k_apply issues together: 4 against a
// three-stage pipeline, KH 128^3 accumulation -6% (FP64) / -27% (FP32 flux); 6 and 8 gained 0.3% more on
// KH at occupancy 3 and cost the ensemble 2%
#ifndef HFV_APPLY_CHUNK
#define HFV_APPLY_CHUNK 4
#endif
#ifndef HFV_GRAD_WAVES
#define HFV_GRAD_WAVES 0
#endif

// Hard errors are recorded for the batch and for the simulation that raised them; a batch keeps
// stepping its other simulations (the failed one stops at its next k_begin).
__device__ __forceinline__ void flush_events(const Events& ev, DevCounters* cnt, SimEngine* eng = nullptr, int s = -1) {
  if (ev.err) {
    atomicCAS(&cnt->err, 0, ev.err);
    if (eng && s >= 0) atomicCAS(&eng[s].err, 0, ev.err);
  }
  if (ev.check_physical) atomicAdd(&cnt->check_physical, (unsigned long long)ev.check_physical);
  if (ev.brent) atomicAdd(&cnt->brent, (unsigned long long)ev.brent);
  if (ev.grad_fallback) atomicAdd(&cnt->grad_fallback, (unsigned long long)ev.grad_fallback);
  if (ev.flux_reset) atomicAdd(&cnt->flux_reset, (unsigned long long)ev.flux_reset);
}

// Block-wide reductions (256 threads = 4 waves); the result is valid in thread 0.
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v) {
  __shared__ uint32_t red_u32[HFV_BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (__lane_id() == 0) red_u32[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red_u32[w];
  return t;
}
__device__ __forceinline__ uint64_t block_min_u64(uint64_t v) {
  __shared__ uint64_t red_u64[HFV_BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  if (__lane_id() == 0) red_u64[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t t = ~0ull;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = red_u64[w] < t ? red_u64[w] : t;
  return t;
}
__device__ __forceinline__ int block_min_i32(int v) {
  __shared__ int red_i32[HFV_BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  if (__lane_id() == 0) red_i32[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0x7fffffff;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = red_i32[w] < t ? red_i32[w] : t;
  return t;
}

// ------------------------------------------------------------------ engine (per simulation)
// Engine::run loop head (engine.rs:152-185): the pending dump sync points are drifted through, then
// the step at ti_next.
__global__ void k_begin(Batch B) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) B.cnt->nfall = 0;
  if (s >= B.nsims) return;
  SimEngine& E = B.eng[s];
  E.stepping = 0;
  E.n_sub = 0;
  E.n_down = 0;
  E.nf_tmp = 0;
  E.n_active = 0;
  E.n_face_active = 0;
  E.ti_min = ~0ull;
  E.min_bin = 127;
  if (E.err) {  // a hard error stops this simulation only; the batch's other simulations go on
    E.force = 0;
    E.done = 1;
    return;
  }
  if (E.force) {  // hfv_step: a step at ti_current with the caller's ti_old
    E.force = 0;
    E.stepping = 1;
  } else {
    if (E.done) return;
    if (!(E.t_current < E.t_end)) {
      E.done = 1;
      return;
    }
    uint64_t cur = E.ti_current;
    int ns = 0;
    for (uint64_t d = E.dump_lo; d < E.dump_hi; d += E.between) {
      if (d > cur) {
        if (ns < HFV_MAX_SUBDRIFTS) E.sub_dt[ns] = (double)(d - cur) * E.time_base;
        ns++;
        cur = d;
      }
      if (E.between == 0) break;
    }
    E.dump_lo = E.dump_hi;
    if (E.ti_next > cur) {
      if (ns < HFV_MAX_SUBDRIFTS) E.sub_dt[ns] = (double)(E.ti_next - cur) * E.time_base;
      ns++;
    }
    if (ns > HFV_MAX_SUBDRIFTS) {
      E.err = E_UNSUPPORTED;
      atomicCAS(&B.cnt->err, 0, E_UNSUPPORTED);
      ns = HFV_MAX_SUBDRIFTS;
    }
    E.n_sub = ns;
    E.ti_current = E.ti_next;
    E.t_current = (double)E.ti_current * E.time_base;
    E.stepping = 1;
  }
  E.dt_pred = (double)(E.ti_current - E.ti_old) * E.time_base;
  E.pred_step = !E.no_pred;
  if (E.pred_step) E.ti_pred += E.ti_current - E.ti_old;
  E.no_pred = 0;
  E.mab = get_max_active_bin(E.ti_current);
}

// Engine::step / queue_step / update_ti_next (engine.rs:194-268).  One block (HFV_END_BLOCK threads) per
// simulation sums the chunk partials of this step (eight loads in flight per thread), then one thread
// moves the engine to its next sync point.
#define HFV_END_BLOCK 1024
__global__ void __launch_bounds__(HFV_END_BLOCK) k_end(Batch B) {
  __shared__ uint32_t sa[HFV_END_BLOCK / 64], sf[HFV_END_BLOCK / 64];
  __shared__ uint64_t st[HFV_END_BLOCK / 64];
  const int s = blockIdx.x;
  SimEngine& E = B.eng[s];
  if (!E.stepping) return;
  const SimDesc& D = B.sims[s];
  const int T = blockDim.x;
  uint32_t na = 0, nfc = 0;
  uint64_t tmin = ~0ull;
  for (int k0 = D.cchunk0 + (int)threadIdx.x; k0 < D.cchunk1; k0 += 8 * T) {
    uint32_t a[8];
    uint64_t t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u * T;
      a[u] = k < D.cchunk1 ? B.P.act[k] : 0u;
      t[u] = k < D.cchunk1 ? B.P.ti[k] : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      na += a[u];
      tmin = t[u] < tmin ? t[u] : tmin;
    }
  }
  for (int k0 = D.fchunk0 + (int)threadIdx.x; k0 < D.fchunk1; k0 += 8 * T) {
    uint32_t a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u * T;
      a[u] = k < D.fchunk1 ? B.P.face[k] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) nfc += a[u];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    na += __shfl_xor(na, off);
    nfc += __shfl_xor(nfc, off);
    uint64_t o = __shfl_xor(tmin, off);
    tmin = o < tmin ? o : tmin;
  }
  if (__lane_id() == 0) {
    sa[threadIdx.x >> 6] = na;
    sf[threadIdx.x >> 6] = nfc;
    st[threadIdx.x >> 6] = tmin;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (E.pred_step) E.chk = 0;  // this step's k_mask has checked the host-written primitives
  na = 0;
  nfc = 0;
  tmin = ~0ull;
  for (int w = 0; w < (T >> 6); ++w) {
    na += sa[w];
    nfc += sf[w];
    tmin = st[w] < tmin ? st[w] : tmin;
  }
  E.n_active = na;
  E.n_face_active = nfc;
  E.ti_min = tmin;
  if (E.err) {  // the reference panics inside this step: it does not complete
    E.done = 1;
    return;
  }
  if (E.n_active == 0 || E.ti_min == ~0ull) {  // the reference panics: "At least one particle must be active"
    E.err = E_UNSUPPORTED;
    atomicCAS(&B.cnt->err, 0, E_UNSUPPORTED);
    E.done = 1;
    return;
  }
  uint64_t ti_next = E.ti_current + E.ti_min;
  E.dump_lo = E.ti_snap;
  if (E.between) {
    while (E.ti_snap < ti_next) E.ti_snap += E.between;
  }
  E.dump_hi = E.ti_snap;
  E.ti_old = E.ti_next;
  E.ti_next = ti_next;
  E.step_count++;
  E.cell_updates += (long long)E.n_active;
  E.face_fluxes += (long long)E.n_face_active;
  E.last_active = (long long)E.n_active;
  E.last_face_active = (long long)E.n_face_active;
}

// ------------------------------------------------------------------ drift (space.rs:833-854)
// Particle::drift + box_wrap / box_reflect (part.rs:107-119, 449-469) for one sub-interval.
// A reflected cell's kept time-extrapolation rate is rewritten from its reflected primitives and
// gradients (its v_rel is not reflected), so that this step's predict (optimal_order.rs:20, after the
// drift) extrapolates the reflected state; FP32-flux contexts first sum their unsummed predict time up
// to the start of this step with the old rate.
__device__ void drift_cell(const Cells& C, const Cfg& cfg, const SimEngine& E, int64_t i, double dt, Events& ev) {
  const int64_t n = C.n;
  const int dim = cfg.dim;
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
        double sh = loc[a] - old[a];
        gc[a] += sh;
        cen[a] += sh;
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
        const St<double> p_old = ld5(C.prim, n, i);
        const St<double> p_new = reflect(p_old, nn);
        st5(C.prim, n, i, p_new);
        // Particle::reflect_gradients (part.rs:390-429), including its use of the already reflected
        // density gradient for the pressure gradient
        const bool f32 = cfg.precision == HFV_FP32_FLUX;
        if (f32) {
          const uint64_t t0 = E.pred_step ? E.ti_pred - (E.ti_current - E.ti_old) : E.ti_pred;
          const double el = (double)(t0 - C.tiex[i]) * E.time_base;
          st_ex(C, i, true, ld_ex(C, i, true) + (p_old.c[0] > 0. ? (-el) * ld_tex(C, i, true) : vacuum<double>()));
          C.tiex[i] = t0;
        }
        Gr<double> g = ld_grad(C, i, f32);
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
        st_grad(C, i, f32, g);
        st_tex(C, i, f32, fr_rate(p_new, round_grad(g, f32), ld3(C.vrel, n, i), cfg.gd));
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
}

// Engine::run's drift through the step's sync points (moving meshes only).
__global__ void __launch_bounds__(HFV_BLOCK) k_drift(Batch B) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const SimEngine& E = B.eng[J.sim];
  const Cfg& cfg = B.sims[J.sim].cfg;
  if (!E.stepping || E.n_sub == 0 || cfg.motion == MOT_FIXED) return;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  if (i >= J.hi) return;
  Events ev;
  for (int k = 0; k < E.n_sub; ++k) drift_cell(B.C, cfg, E, i, E.sub_dt[k], ev);
  flush_events(ev, B.cnt, B.eng, J.sim);
}

// ------------------------------------------------------------------ k_mask
// Active mask (optimal_order.rs:17) and predict (euler_eqns_fvs.rs:20-24, part.rs:121-149) of every
// cell.  mode 0: step; 1: predict only (dt from the engine's dt_pred); 2: mask only.
// FP32-flux contexts defer the step's predict (k_begin advanced ti_pred; the flux kernel's readers add
// rate x elapsed time, face_side): check_physical runs only after host writes of the primitives (a no-op otherwise,
// since c2p and the apply_flux reset leave physical states), settling the sums first.  Mode 1 settles
// the sums and adds dt_pred at once.
__global__ void __launch_bounds__(HFV_BLOCK) k_mask(Batch B, int mode) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimEngine& E = B.eng[s];
  if (mode != 1 && !E.stepping) return;
  const Cfg& cfg = B.sims[s].cfg;
  const Cells& C = B.C;
  const int64_t n = C.n;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  const bool valid = i < J.hi;
  bool act = false;
  Events ev;
  if (valid) {
    if (mode != 1) {
      act = C.bin[i] <= E.mab;
      C.act[i] = act;
    }
    const bool f32 = cfg.precision == HFV_FP32_FLUX;
    if (mode != 2 && (!f32 || mode == 1 || E.chk)) {
      St<double> prim = ld5(C.prim, n, i);
      // the rate of this prim/grad/v_rel, kept by k_grad (k_pack_fr after host changes);
      // extrapolations += time_extrapolations(dt) (its vacuum branch adds +0)
      const St<double> rate = prim.c[0] > 0. ? ld_tex(C, i, f32) : vacuum<double>();
      if (f32) {
        const double dt = (double)(E.ti_pred - C.tiex[i]) * E.time_base + (mode == 1 ? E.dt_pred : 0.);
        st_ex(C, i, true, ld_ex(C, i, true) + (prim.c[0] > 0. ? (-dt) * rate : vacuum<double>()));
        C.tiex[i] = E.ti_pred;
      } else {
        St<double> te = prim.c[0] > 0. ? (-E.dt_pred) * rate : vacuum<double>();
        st_ex(C, i, false, ld_ex(C, i, false) + te);
      }
      check_physical(prim, ev);
      if (ev.check_physical) {
        st5(C.prim, n, i, prim);
        store_pc(C.pc, i, prim, ld3(C.cen, n, i));
        if (B.sims[s].staged) {
          // the flux record holds this rate already; a reset changes prim and the sound speed
          double* r = C.fr + HFV_FR_W * (uint64_t)i;
          if (f32) {
            float* w = (float*)r;
#pragma unroll
            for (int q = 0; q < 5; ++q) w[HFV_FR32_PRIM + q] = (float)prim.c[q];
            w[HFV_FR32_CS] = (float)fr_cs(prim, cfg.gd);
          } else {
#pragma unroll
            for (int q = 0; q < 5; ++q) r[HFV_FR_PRIM + q] = prim.c[q];
            r[HFV_FR_CS] = fr_cs(prim, cfg.gd);
          }
        }
      }
      // volume += dvdt * dt clamped to [V/2, 2V]: dvdt is never set on the path (part.rs:130-131)
    }
  }
  if (mode != 1) {
    uint32_t cnt = block_sum_u32(act ? 1u : 0u);
    if (threadIdx.x == 0) {
      B.P.act[J.chunk] = cnt;
    }
  }
  if (valid) flush_events(ev, B.cnt, B.eng, s);
}

// ------------------------------------------------------------------ 1D moving mesh on the device
// Space::volume_calculation for 1D (space.rs:236-263; semantics pinned by test_init_1d,
// space.rs:1453-1491): generators keep their order along x, so the face topology is fixed; face k of a
// simulation sits between its cells k-1 and k; walls are faces 0 and n (non periodic).  On a periodic
// line the order is cyclic: a generator that drifted through the box edge keeps its neighbours, and the one
// face where the order steps down (face 0 until something wraps) joins its left generator to the image of
// its right one a box to the right, carrying the periodic shift.
__global__ void __launch_bounds__(HFV_BLOCK) k_mesh1d_faces(Batch B) {
  const BlockJob J = B.fjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.moving1d || !B.eng[s].stepping) return;
  const int64_t f = (int64_t)J.lo + threadIdx.x;
  if (f >= J.hi) return;
  int64_t k = f - D.face_base, n = D.n, cb = D.cell_base;
  double x;
  if (!D.periodic) {
    if (k >= 1 && k < n) {
      double a = B.C.loc[cb + k - 1], b = B.C.loc[cb + k];
      if (!(a < b)) {
        atomicCAS(&B.cnt->err, 0, E_UNSUPPORTED);
        atomicCAS(&B.eng[s].err, 0, E_UNSUPPORTED);
      }
      x = 0.5 * (a + b);
    } else {
      x = (k == 0) ? 0. : D.cfg.box[0];
    }
  } else {
    const double a = B.C.loc[cb + (k == 0 ? n - 1 : k - 1)], b = B.C.loc[cb + k];
    const bool down = !(a < b);
    x = down ? 0.5 * (a + (b + D.cfg.box[0])) : 0.5 * (a + b);
    B.F.shc[f] = down ? 1 : 0;  // +box along x
    if (down) atomicAdd(&B.eng[s].n_down, 1);
  }
  B.F.fcen[f] = x;
}
__global__ void __launch_bounds__(HFV_BLOCK) k_mesh1d_cells(Batch B) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const SimDesc& D = B.sims[J.sim];
  if (!D.moving1d || !B.eng[J.sim].stepping) return;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  if (i >= J.hi || !B.C.act[i]) return;
  const int64_t N = B.C.n;
  int64_t k = i - D.cell_base, n = D.n, fb = D.face_base;
  double xl, xr;
  if (!D.periodic) {
    xl = B.F.fcen[fb + k];
    xr = B.F.fcen[fb + k + 1];
  } else {
    if (B.eng[J.sim].n_down != 1) {  // not a rotation of the line's order: the topology changed
      atomicCAS(&B.cnt->err, 0, E_UNSUPPORTED);
      atomicCAS(&B.eng[J.sim].err, 0, E_UNSUPPORTED);
    }
    // the left face's centroid is in cell k-1's frame: a box to the right when it carries the shift
    xl = B.F.fcen[fb + k] - (B.F.shc[fb + k] ? D.cfg.box[0] : 0.);
    xr = (k == n - 1) ? B.F.fcen[fb] : B.F.fcen[fb + k + 1];
  }
  double V = xr - xl;
  B.C.vol[i] = V;
  B.C.cen[i] = 0.5 * (xl + xr);
  B.C.cen[N + i] = 0.;
  B.C.cen[2 * N + i] = 0.;
  B.C.srad[i] = 1.5 * V;
}

// ------------------------------------------------------------------ k_flux
// FiniteVolumeSolver::compute_fluxes: gate (euler_eqns_fvs.rs:26-48) + flux_exchange (flux.rs:23-100)
// + flux_exchange_boundary (flux.rs:102-195).  T: arithmetic of reconstruction and solve; SOLVER >= 0
// specialises the kernel for one Riemann solver (-1: runtime switch).
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

// Per-cell inputs of the flux phase gathered from the cell fields.  The sound speed is k_mask's
// expression of the same primitives (a check_physical reset in k_mask leaves rho = 0, where neither the
// sound speed nor the rate is used).  With T = float the fields that the FP32 flux
// record holds in FP32 (primitives, sound speed, gc) are rounded the same way, so that gathered and
// staged faces compute the same bits.
template <class T>
__device__ __forceinline__ double rnd(double x) {
  return (double)(T)x;
}
template <class T>
struct SoaCell {
  const Cells* C;
  const Gas<double>* g;
  const SimEngine* E;
  uint32_t i;
  __device__ __forceinline__ SoaCell(const Cells& c, const Gas<double>& gd, const SimEngine& e, int32_t idx)
      : C(&c), g(&gd), E(&e), i((uint32_t)idx) {}
  __device__ __forceinline__ double dt() const { return ldo(C->dt, i); }
  // predict time not yet summed into the stored extrapolations (FP32-flux contexts)
  __device__ __forceinline__ double el() const { return (double)(E->ti_pred - C->tiex[i]) * E->time_base; }
  __device__ __forceinline__ St<double> prim() const {
    const St<double> p = ld5(C->prim, C->n, i);
    return mkst(rnd<T>(p.c[0]), mk3(rnd<T>(p.c[1]), rnd<T>(p.c[2]), rnd<T>(p.c[3])), rnd<T>(p.c[4]));
  }
  __device__ __forceinline__ double rho() const { return rnd<T>(ldo(C->prim, i)); }
  __device__ __forceinline__ V3<double> pvel() const {
    return mk3(rnd<T>(ldo(C->prim + C->n, i)), rnd<T>(ldo(C->prim + 2 * C->n, i)), rnd<T>(ldo(C->prim + 3 * C->n, i)));
  }
  __device__ __forceinline__ double cs() const { return rnd<T>(fr_cs(ld5(C->prim, C->n, i), *g)); }
  __device__ __forceinline__ V3<double> grad(int q) const {
    if constexpr (sizeof(T) == 4) {
      const float* a = (const float*)C->grad;
      const uint64_t n = (uint64_t)C->n;
      return mk3((double)a[(3 * q) * n + i], (double)a[(3 * q + 1) * n + i], (double)a[(3 * q + 2) * n + i]);
    }
    return mk3(ldo(C->grad + (3 * q) * C->n, i), ldo(C->grad + (3 * q + 1) * C->n, i), ldo(C->grad + (3 * q + 2) * C->n, i));
  }
  __device__ __forceinline__ double ex(int q) const {
    if constexpr (sizeof(T) == 4) return ((const float*)C->extrap)[HFV_EX_W * (uint64_t)i + q];
    return C->extrap[HFV_EX_W * (uint64_t)i + q];
  }
  __device__ __forceinline__ St<double> rate() const { return ld_tex(*C, i, sizeof(T) == 4); }
  __device__ __forceinline__ V3<double> gc() const {
    const V3<double> v = ld3(C->gc, C->n, i);
    return mk3(rnd<T>(v.x), rnd<T>(v.y), rnd<T>(v.z));
  }
  __device__ __forceinline__ V3<double> cen() const { return ld3(C->cen, C->n, i); }
  __device__ __forceinline__ V3<double> loc() const { return ld3(C->loc, C->n, i); }
  __device__ __forceinline__ V3<double> v() const { return ld3(C->v, C->n, i); }
};
// Per-cell inputs of the flux phase from a cell's flux record and extrapolations staged in LDS
// (static geometry): FP64 record fields at HFV_FR_*, extrapolations at HFV_FR_EX.
struct RecCell {
  const double* r;
  const Cells* C;
  uint32_t i;
  __device__ __forceinline__ RecCell(const double* rec, const Cells& c, int32_t idx) : r(rec), C(&c), i((uint32_t)idx) {}
  __device__ __forceinline__ double dt() const { return r[HFV_FR_DT]; }
  __device__ __forceinline__ St<double> prim() const {
    return mkst(r[HFV_FR_PRIM], mk3(r[HFV_FR_PRIM + 1], r[HFV_FR_PRIM + 2], r[HFV_FR_PRIM + 3]), r[HFV_FR_PRIM + 4]);
  }
  __device__ __forceinline__ double rho() const { return r[HFV_FR_PRIM]; }
  __device__ __forceinline__ V3<double> pvel() const { return mk3(r[HFV_FR_PRIM + 1], r[HFV_FR_PRIM + 2], r[HFV_FR_PRIM + 3]); }
  __device__ __forceinline__ double cs() const { return r[HFV_FR_CS]; }
  __device__ __forceinline__ V3<double> grad(int q) const {
    return mk3(r[HFV_FR_GRAD + 3 * q], r[HFV_FR_GRAD + 3 * q + 1], r[HFV_FR_GRAD + 3 * q + 2]);
  }
  __device__ __forceinline__ double ex(int q) const { return r[HFV_FR_EX + q]; }
  __device__ __forceinline__ double el() const { return 0.; }  // FP64 contexts sum every step
  __device__ __forceinline__ St<double> rate() const {
    return mkst(r[HFV_FR_TEX], mk3(r[HFV_FR_TEX + 1], r[HFV_FR_TEX + 2], r[HFV_FR_TEX + 3]), r[HFV_FR_TEX + 4]);
  }
  __device__ __forceinline__ V3<double> gc() const { return mk3(r[HFV_FR_GC], r[HFV_FR_GC + 1], r[HFV_FR_GC + 2]); }
  __device__ __forceinline__ V3<double> cen() const { return ld3(C->cen, C->n, i); }
  __device__ __forceinline__ V3<double> loc() const { return ld3(C->loc, C->n, i); }
  __device__ __forceinline__ V3<double> v() const { return ld3(C->v, C->n, i); }
};
// The same from an FP32 record staged in LDS (words at HFV_FR32_*, extrapolations at HFV_FR32_EX).
struct RecCellF {
  const float* r;
  const Cells* C;
  uint32_t i;
  __device__ __forceinline__ RecCellF(const float* rec, const Cells& c, int32_t idx) : r(rec), C(&c), i((uint32_t)idx) {}
  __device__ __forceinline__ double dt() const { return __hiloint2double(__float_as_int(r[1]), __float_as_int(r[0])); }
  __device__ __forceinline__ St<double> prim() const {
    return mkst((double)r[HFV_FR32_PRIM], mk3((double)r[HFV_FR32_PRIM + 1], (double)r[HFV_FR32_PRIM + 2],
                (double)r[HFV_FR32_PRIM + 3]), (double)r[HFV_FR32_PRIM + 4]);
  }
  __device__ __forceinline__ double rho() const { return r[HFV_FR32_PRIM]; }
  __device__ __forceinline__ V3<double> pvel() const {
    return mk3((double)r[HFV_FR32_PRIM + 1], (double)r[HFV_FR32_PRIM + 2], (double)r[HFV_FR32_PRIM + 3]);
  }
  __device__ __forceinline__ double cs() const { return r[HFV_FR32_CS]; }
  __device__ __forceinline__ V3<double> grad(int q) const {
    return mk3((double)r[HFV_FR32_GRAD + 3 * q], (double)r[HFV_FR32_GRAD + 3 * q + 1], (double)r[HFV_FR32_GRAD + 3 * q + 2]);
  }
  __device__ __forceinline__ double ex(int q) const { return r[HFV_FR32_EX + q]; }
  __device__ __forceinline__ double el() const { return r[HFV_FR32_EL]; }
  __device__ __forceinline__ St<double> rate() const {
    return mkst((double)r[HFV_FR32_TEX], mk3((double)r[HFV_FR32_TEX + 1], (double)r[HFV_FR32_TEX + 2],
                (double)r[HFV_FR32_TEX + 3]), (double)r[HFV_FR32_TEX + 4]);
  }
  __device__ __forceinline__ V3<double> gc() const {
    return mk3((double)r[HFV_FR32_GC], (double)r[HFV_FR32_GC + 1], (double)r[HFV_FR32_GC + 2]);
  }
  __device__ __forceinline__ V3<double> cen() const { return ld3(C->cen, C->n, i); }
  __device__ __forceinline__ V3<double> loc() const { return ld3(C->loc, C->n, i); }
  __device__ __forceinline__ V3<double> v() const { return ld3(C->v, C->n, i); }
};
// One side's reconstructed state at the face before limiting (flux.rs:52-73):
// ((W + grad . dx) + extrapolations) + time_extrapolations(-dt/2), in T.  FP32-flux contexts keep the
// steps' predicts since the cell's tag unsummed (el): the rate is constant over them, so the
// extrapolations' pending part and the half step back (dte = -dt/2) share one product, rate x -(dte + el).
template <class T, class Acc>
__device__ __forceinline__ St<T> face_side(const Acc& X, const St<T>& pT, V3<T> dxT, T dte) {
  St<T> out;
#pragma unroll
  for (int q = 0; q < 5; ++q) out.c[q] = (pT.c[q] + dot(cvt3<T>(X.grad(q)), dxT)) + (T)X.ex(q);
  if (pT.c[0] > T(0)) {
    const St<T> rate = cvtst<T>(X.rate());
    if constexpr (sizeof(T) == 4) out = out + (-(dte + (T)X.el())) * rate;
    else out = out + (-dte) * rate;
  } else {
    out = out + vacuum<T>();
  }
  return out;
}

// The face's record (dt*F, v_max, a_over_r) when computed; false when gated off (record zero).
// Interior and wall faces build the face-frame problem (states, face velocity, normal, dt, area) in
// their own branch and share one Riemann call, which is inlined (no argument traffic through scratch).
// STATIC (fixed mesh, zero particle velocities): the centroid offset and a_over_r come from the face's
// precomputed line (same expressions, evaluated once by k_geom_faces) and the face velocity is the
// value the formula yields for v = 0, namely +0 in every component.
template <class T, int SOLVER, bool DEFER, bool STATIC, class Acc>
__device__ __forceinline__ bool face_flux(const Batch& B, const Cfg& cfg, int64_t f, const Acc& L, const Acc& R,
                                          int32_t r, bool la, bool ra, double* outv, Events& ev) {
  const Faces& F = B.F;
  const int64_t nf = F.nf;
  const Gas<T>& g = gas_for<T>(cfg);
  SolverCfg sc = cfg.solver;
  if (SOLVER >= 0) sc.kind = SOLVER;
#pragma unroll
  for (int k = 0; k < 7; ++k) outv[k] = 0.;
  const double dtl = L.dt();
  St<T> WL, WR;  // face-frame states
  V3<T> vfT;     // face velocity
  double dt, v_max = 0., len_dx = 1., a_over_r = 0.;
  if (r >= 0) {
    // ---- gate
    const double dtr = R.dt();
    if ((!la && !ra) || (dtr < dtl && !ra) || (dtl < dtr && !la)) return false;
    dt = hmin(dtl, dtr);
    HFV_FSTAGE();
    // ---- geometry and the two reconstructed sides
    const V3<double> fc = ld3(F.fcen, nf, f);
    const V3<double> sh = decode_shift(F.shc[f], cfg.box);
    const T dte = (T)(-0.5 * dt);
    const V3<double> dx_left = fc - L.gc();
    const St<T> plT = cvtst<T>(L.prim());
    const St<T> left_dash = face_side<T>(L, plT, cvt3<T>(dx_left), dte);
    HFV_FSTAGE();
    const V3<double> dx_right = (fc - R.gc()) - sh;
    const St<T> prT = cvtst<T>(R.prim());
    const St<T> right_dash = face_side<T>(R, prT, cvt3<T>(dx_right), dte);
    HFV_FSTAGE();
    // ---- pairwise limiter (gradients.rs:105-152)
    const T dxg = (T)len(dx_left - dx_right);
    WL = pairwise_limiter(plT, prT, left_dash, fdiv(len(cvt3<T>(dx_left)), dxg), ev);
    WR = pairwise_limiter(prT, plT, right_dash, fdiv(len(cvt3<T>(dx_right)), dxg), ev);
    HFV_FSTAGE();
    // ---- signal speed (flux.rs:43-50), face velocity (flux.rs:75-78), a_over_r: FP64
    if (STATIC) {
      const double4 st = *(const double4*)(F.fs + 4 * (uint64_t)f);
      if (L.rho() > 0.) v_max += L.cs();
      if (R.rho() > 0.) v_max += R.cs();
      v_max -= hmin(dot(R.pvel() - L.pvel(), mk3(st.x, st.y, st.z)), 0.);
      a_over_r = st.w;
      vfT = mk3(T(0), T(0), T(0));
    } else {
      {
        const V3<double> dx_centroid = (R.cen() + sh) - L.cen();
        if (L.rho() > 0.) v_max += L.cs();
        if (R.rho() > 0.) v_max += R.cs();
        v_max -= hmin(dot(R.pvel() - L.pvel(), dx_centroid), 0.);
      }
      HFV_FSTAGE();
      {
        const V3<double> locl = L.loc(), locr = R.loc();
        const V3<double> dx = (locr + sh) - locl;
        const V3<double> vl = L.v(), vr = R.v();
        const V3<double> midpoint = 0.5 * ((locl + locr) + sh);
        const double fac = dot(vr - vl, fc - midpoint) / len2(dx);
        vfT = cvt3<T>((0.5 * (vl + vr)) - (fac * dx));
        len_dx = len(dx);
      }
    }
    HFV_FSTAGE();
    WL = boost(WL, -vfT);
    WR = boost(WR, -vfT);
  } else {
    // ---- wall face (flux_exchange_boundary, flux.rs:102-195)
    if (!la) return false;
    const V3<double> fc = ld3(F.fcen, nf, f);
    const V3<double> nrm = ld3(F.nrm, nf, f);
    const V3<double> gc = L.gc();
    // Particle::reflect (part.rs:374-383)
    const V3<double> gc_r = gc + ((2. * dot(fc - gc, nrm)) * nrm);
    const V3<double> dx_face = fc - gc;
    const double dx_gradients = len(gc_r - gc);
    V3<double> dx_centroid;
    if (STATIC) {
      const double4 st = *(const double4*)(F.fs + 4 * (uint64_t)f);
      dx_centroid = mk3(st.x, st.y, st.z);
      a_over_r = st.w;
    } else {
      const V3<double> loc = L.loc(), cen = L.cen();
      const V3<double> loc_r = loc + ((2. * dot(fc - loc, nrm)) * nrm);
      const V3<double> cen_r = cen + ((2. * dot(fc - cen, nrm)) * nrm);
      dx_centroid = cen_r - cen;
      len_dx = len(loc_r - loc);
    }
    const St<double> p = L.prim();
    dt = dtl;
    if (p.c[0] > 0.) v_max += L.cs();
    const St<T> pT = cvtst<T>(p);
    const V3<T> dxfT = cvt3<T>(dx_face);
    const V3<T> nT = cvt3<T>(nrm);
    St<T> dash = face_side<T>(L, pT, dxfT, (T)(-0.5 * dt));
    const T dx_fac = fdiv(len(dxfT), (T)dx_gradients);
    if (cfg.boundary == BND_REFLECTIVE) {
      St<T> rp = reflect(pT, nT);
      dash = pairwise_limiter(pT, rp, dash, dx_fac, ev);
      St<double> rpd = reflect(p, nrm);
      v_max -= hmin(dot(vel(rpd) - vel(p), dx_centroid), 0.);
      WR = reflect(dash, nT);
    } else if (cfg.boundary == BND_OPEN) {
      dash = pairwise_limiter(pT, pT, dash, dx_fac, ev);
      WR = dash;
    } else if (cfg.boundary == BND_VACUUM) {
      St<T> vac = vacuum<T>();
      dash = pairwise_limiter(pT, vac, dash, dx_fac, ev);
      v_max += hmin(dot(vel(p), dx_centroid), 0.);
      WR = vac;
    } else {
      ev.fail(E_UNSUPPORTED);
      WR = dash;
    }
    WL = dash;
    vfT = mk3(T(0), T(0), T(0));
  }
  // ---- Riemann problem in the face frame (the wall problem has no boost: v_face = 0)
  const V3<double> nrm = ld3(F.nrm, nf, f);
  const double area = F.area[f];
  const St<T> Fl = riemann_flux<T, DEFER>(sc, WL, WR, vfT, cvt3<T>(nrm), g, cfg.gd, ev);
  const St<double> rec = dt * (area * cvtst<double>(Fl));
#pragma unroll
  for (int k = 0; k < 5; ++k) outv[k] = rec.c[k];
  outv[5] = v_max;
  outv[6] = STATIC ? a_over_r : area / len_dx;
  return true;
}

// Face record: FP64 contexts 8 doubles (64 bytes) per face; FP32-flux contexts 8 floats (32 bytes):
// dt*F[5], v_max, a_over_r, unused.  The FP32 record rounds dt*A*F once; both cells accumulate that
// same value in FP64 (with opposite signs), so conservation stays exact.
__device__ __forceinline__ void store_rec(double* rec, int64_t f, const double* outv, bool f32) {
  if (f32) {
    float4* dst = (float4*)((float*)rec + 8 * (uint64_t)f);
    dst[0] = make_float4((float)outv[0], (float)outv[1], (float)outv[2], (float)outv[3]);
    dst[1] = make_float4((float)outv[4], (float)outv[5], (float)outv[6], 0.f);
    return;
  }
  double2* dst = (double2*)(rec + 8 * (uint64_t)f);
#pragma unroll
  for (int q = 0; q < 3; ++q) dst[q] = make_double2(outv[2 * q], outv[2 * q + 1]);
  dst[3] = make_double2(outv[6], 0.);
}

// Faces with an active cell get their record (zero when gated off) and cur = 1; the others cur = 0.
// With SOLVER == AIRS, faces that need the exact solver are queued for k_flux_fallback.
// STATIC: when the chunk has a staging list and a face with an active cell, the block stages its
// distinct cells' flux records and extrapolations in LDS by whole lines (each cell is read once per
// block, instead of every face gathering its two cells field by field); the list, then every record
// load of a thread, are issued together (two memory round trips).  Chunks without a list gather.
template <class T, int SOLVER, bool STATIC>
__device__ __forceinline__ void flux_chunk(const Batch& B, const BlockJob J) {
  if (J.sim < 0) return;
  const int s = J.sim;
  if (!B.eng[s].stepping || B.sims[s].static_geom != (STATIC ? 1 : 0)) return;
  const Cfg& cfg = B.sims[s].cfg;
  const int64_t f = (int64_t)J.lo + threadIdx.x;
  constexpr bool DEFER = SOLVER == SOLVER_AIRS;
  // LDS: the staged records, reused for the block's output records (written out as whole lines)
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int STAGE_D = !STATIC ? 0 : F32 ? (HFV_STAGE_MAX * HFV_FR32_LDS + 1) / 2 : HFV_STAGE_MAX * HFV_FR_LDS;
  constexpr int REC_D = HFV_BLOCK * 8;
  __shared__ double lds[STAGE_D > REC_D ? STAGE_D : REC_D];
  __shared__ uint8_t curs[HFV_BLOCK];
  __shared__ int32_t scl[STATIC ? HFV_STAGE_MAX : 1];
  double2* recs = (double2*)lds;
  const int ns = STATIC && B.sims[s].staged ? (int)B.F.scnt[J.chunk] : 0;  // records kept only when staged
  bool computed = false;
  bool cur = false;
  Events ev;
  double outv[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) outv[k] = 0.;
  int32_t l = 0, r = -1;
  bool la = false, ra = false;
  int32_t myc = 0;
  if (STATIC && (int)threadIdx.x < ns) myc = B.F.scell[(uint64_t)J.chunk * HFV_STAGE_MAX + threadIdx.x];
  if (f < J.hi) {
    l = B.F.left[f];
    r = B.F.right[f];
    if (!STATIC && l < 0) {  // past a device-rebuilt mesh's faces (dynamic geometry only): no cells
      l = 0;
      r = -1;
    } else {
      la = B.C.act[l];
      ra = r >= 0 && B.C.act[r];
      cur = la || ra;
    }
    B.F.cur[f] = cur ? 1 : 0;
  }
  if (STATIC) {
    if ((int)threadIdx.x < ns) scl[threadIdx.x] = myc;
    const int ncur = __syncthreads_count(cur);  // also publishes scl
    if (ns > 0 && ncur > 0) {  // block-uniform
      constexpr int RD2 = F32 ? 8 : HFV_FR_W / 2;  // double2 per record
      constexpr int NR = HFV_STAGE_MAX * RD2 / HFV_BLOCK;
      constexpr int NE = (HFV_STAGE_MAX * HFV_EX_W + HFV_BLOCK - 1) / HFV_BLOCK;
      static_assert(HFV_STAGE_MAX * RD2 % HFV_BLOCK == 0, "record slots per thread");
      double2 rv[NR];
      double xv[NE];
      uint64_t tv = 0;  // FP32: the staged cell's extrapolation tag
      if (F32 && (int)threadIdx.x < ns) tv = B.C.tiex[(uint32_t)myc];
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const int k = (int)threadIdx.x + j * HFV_BLOCK, c = k / RD2, q = k % RD2;
        if (c < ns && (F32 || q < HFV_FR_USED / 2))
          rv[j] = ((const double2*)(B.C.fr + HFV_FR_W * (uint64_t)(uint32_t)scl[c]))[q];
      }
#pragma unroll
      for (int j = 0; j < NE; ++j) {
        const int k = (int)threadIdx.x + j * HFV_BLOCK, c = k / HFV_EX_W, q = k - HFV_EX_W * c;
        if (c < ns)
          xv[j] = F32 ? (double)((const float*)B.C.extrap)[HFV_EX_W * (uint64_t)(uint32_t)scl[c] + q]
                      : B.C.extrap[HFV_EX_W * (uint64_t)(uint32_t)scl[c] + q];
      }
      if (F32) {
        float* ldsf = (float*)lds;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int k = (int)threadIdx.x + j * HFV_BLOCK, c = k / RD2, q = k % RD2;
          if (c < ns) {
            const int2 a = *(const int2*)&rv[j].x, b = *(const int2*)&rv[j].y;
            float* d = ldsf + c * HFV_FR32_LDS + 4 * q;
            d[0] = __int_as_float(a.x);
            d[1] = __int_as_float(a.y);
            d[2] = __int_as_float(b.x);
            if (4 * q + 3 < HFV_FR32_USED) d[3] = __int_as_float(b.y);
          }
        }
#pragma unroll
        for (int j = 0; j < NE; ++j) {
          const int k = (int)threadIdx.x + j * HFV_BLOCK, c = k / HFV_EX_W, q = k - HFV_EX_W * c;
          if (c < ns) ldsf[c * HFV_FR32_LDS + HFV_FR32_EX + q] = (float)xv[j];
        }
        if ((int)threadIdx.x < ns) {
          const SimEngine& E = B.eng[s];
          ldsf[threadIdx.x * HFV_FR32_LDS + HFV_FR32_EL] = (float)((double)(E.ti_pred - tv) * E.time_base);
        }
      } else {
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int k = (int)threadIdx.x + j * HFV_BLOCK, c = k / RD2, q = k % RD2;
          if (c < ns && q < HFV_FR_USED / 2) {
            lds[c * HFV_FR_LDS + 2 * q] = rv[j].x;
            lds[c * HFV_FR_LDS + 2 * q + 1] = rv[j].y;
          }
        }
#pragma unroll
        for (int j = 0; j < NE; ++j) {
          const int k = (int)threadIdx.x + j * HFV_BLOCK, c = k / HFV_EX_W, q = k - HFV_EX_W * c;
          if (c < ns) lds[c * HFV_FR_LDS + HFV_FR_EX + q] = xv[j];
        }
      }
      __syncthreads();
      if (cur) {
        const uchar2 fl = *(const uchar2*)(B.F.floc + 2 * (uint64_t)f);
        const int rs = r >= 0 ? fl.y : fl.x;
        if constexpr (F32) {
          const float* ldsf = (const float*)lds;
          const RecCellF L(ldsf + fl.x * HFV_FR32_LDS, B.C, l), R(ldsf + rs * HFV_FR32_LDS, B.C, r);
          computed = face_flux<T, SOLVER, DEFER, true, RecCellF>(B, cfg, f, L, R, r, la, ra, outv, ev);
        } else {
          const RecCell L(lds + fl.x * HFV_FR_LDS, B.C, l), R(lds + rs * HFV_FR_LDS, B.C, r);
          computed = face_flux<T, SOLVER, DEFER, true, RecCell>(B, cfg, f, L, R, r, la, ra, outv, ev);
        }
      }
      __syncthreads();  // the staged records are dead from here on
    } else if (cur) {
      computed = face_flux<T, SOLVER, DEFER, true, SoaCell<T>>(B, cfg, f, SoaCell<T>(B.C, cfg.gd, B.eng[s], l),
                                                               SoaCell<T>(B.C, cfg.gd, B.eng[s], r), r, la, ra, outv, ev);
    }
  } else if (cur) {
    computed = face_flux<T, SOLVER, DEFER, false, SoaCell<T>>(B, cfg, f, SoaCell<T>(B.C, cfg.gd, B.eng[s], l),
                                                              SoaCell<T>(B.C, cfg.gd, B.eng[s], r), r, la, ra, outv, ev);
  }
  // records of faces with an active cell (deferred ones zero here, overwritten by k_flux_fallback)
  curs[threadIdx.x] = cur ? 1 : 0;
  if constexpr (F32) {
    float4* rf = (float4*)lds;
    rf[2 * threadIdx.x] = make_float4((float)outv[0], (float)outv[1], (float)outv[2], (float)outv[3]);
    rf[2 * threadIdx.x + 1] = make_float4((float)outv[4], (float)outv[5], (float)outv[6], 0.f);
    __syncthreads();
    float4* dst = (float4*)((float*)B.F.rec + 8 * (uint64_t)J.lo);
    const int nvalid = 2 * (J.hi - J.lo);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = q * HFV_BLOCK + (int)threadIdx.x;
      if (c < nvalid && curs[c >> 1]) dst[c] = rf[c];
    }
  } else {
    recs[4 * threadIdx.x + 0] = make_double2(outv[0], outv[1]);
    recs[4 * threadIdx.x + 1] = make_double2(outv[2], outv[3]);
    recs[4 * threadIdx.x + 2] = make_double2(outv[4], outv[5]);
    recs[4 * threadIdx.x + 3] = make_double2(outv[6], 0.);
    __syncthreads();
    double2* dst = (double2*)(B.F.rec + 8 * (uint64_t)J.lo);
    const int nvalid = 4 * (J.hi - J.lo);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = q * HFV_BLOCK + (int)threadIdx.x;
      if (c < nvalid && curs[c >> 2]) dst[c] = recs[c];
    }
  }
  // deferred faces: wave-aggregated append (rare)
  const unsigned long long m = __ballot(ev.deferred != 0);
  if (m) {
    const int leader = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (__lane_id() == leader) base = atomicAdd(&B.cnt->nfall, (unsigned int)__popcll(m));
    base = __shfl(base, leader);
    if (ev.deferred) B.fall[base + __popcll(m & ((1ull << __lane_id()) - 1ull))] = make_int2((int)f, s);
  }
  uint32_t cnt = block_sum_u32(computed ? 1u : 0u);
  if (threadIdx.x == 0) B.P.face[J.chunk] = cnt;
  ev.deferred = 0;
  flush_events(ev, B.cnt, B.eng, s);
}

// One block per fpt face chunks (fpt faces per thread: 1, 2 or 4, the autotuner's launch-shape choice):
// block b takes the table's positions b, b + grid, .. (the chunks of a block keep the XCD their positions
// were placed on whenever the grid is a multiple of 8).
template <class T, int SOLVER, bool STATIC, int FPT>
__global__ void __launch_bounds__(HFV_BLOCK, !STATIC || sizeof(T) != 4 ? HFV_FLUX_WAVES
                                             : SOLVER == SOLVER_AIRS ? HFV_FLUX_WAVES_F32_AIRS : HFV_FLUX_WAVES_F32)
    k_flux(Batch B) {
#pragma unroll
  for (int q = 0; q < FPT; ++q) {
    flux_chunk<T, SOLVER, STATIC>(B, B.fjobs[blockIdx.x + q * gridDim.x]);
    if (q + 1 < FPT) __syncthreads();  // the chunk's LDS is dead before the next chunk stages
  }
}

// The queued faces of this step with the full AIRS solver (exact fallback), grid-stride.
template <class T>
__global__ void __launch_bounds__(HFV_BLOCK) k_flux_fallback(Batch B) {
  const unsigned int count = B.cnt->nfall;
  for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
    const int2 e = B.fall[k];
    const int64_t f = e.x;
    const int s = e.y;
    const int32_t l = B.F.left[f], r = B.F.right[f];
    const bool la = B.C.act[l], ra = r >= 0 && B.C.act[r];
    Events ev;
    double outv[7];
    const Cfg& cfg = B.sims[s].cfg;
    face_flux<T, SOLVER_AIRS, false, false, SoaCell<T>>(B, cfg, f, SoaCell<T>(B.C, cfg.gd, B.eng[s], l),
                                                        SoaCell<T>(B.C, cfg.gd, B.eng[s], r), r, la, ra, outv, ev);
    store_rec(B.F.rec, f, outv, sizeof(T) == 4);
    flush_events(ev, B.cnt, B.eng, s);
  }
}

// ------------------------------------------------------------------ k_apply
// runner apply_fluxes (hydrodynamics.rs:13-39) + kick2 (:51-57) + convert_conserved_to_primitive
// (euler_eqns_fvs.rs:63-69, part.rs:219-261).  One thread per cell walks its faces in CSR order, so the
// FP64 accumulation order is the reference's and needs no atomics.  Faces without an active cell are
// gated off in the reference (zero records); they are skipped here.
// mode bit 1: accumulate + apply + kick2; mode bit 2: conserved -> primitive.
__device__ __forceinline__ St<double> c2p(const St<double>& c, double vol, const Gas<double>& g, Events& ev) {
  St<double> prim;
  if (!(c.c[0] >= 0.)) ev.fail(E_NEG_MASS);
  if (!(c.c[4] >= 0.)) ev.fail(E_NEG_ENERGY);
  if (c.c[0] == 0.) {
    if (c.c[4] != 0.) ev.fail(E_NEG_ENERGY);
    prim = vacuum<double>();
  } else if (c.c[0] > 0.) {
    double m_inv = 1. / c.c[0];
    double density = c.c[0] / vol;
    V3<double> velocity = m_inv * vel(c);
    double thermal = c.c[4] - 0.5 * len2(vel(c)) * m_inv;
    double u = thermal * m_inv;
    double pressure = gas_p_from_u(g, u, density);
    if (!(density >= 0.) || !(pressure >= 0.)) ev.fail(E_NEG_PRESSURE);
    prim = mkst(density, velocity, pressure);
  } else {
    prim = vacuum<double>();
  }
#pragma unroll
  for (int q = 0; q < 5; ++q)
    if (!hfinite(prim.c[q])) ev.fail(E_NONFINITE);
  return prim;
}


// mode bit 1: accumulate + apply + kick2; bit 2: conserved -> primitive; bit 4: update the pc pack
// (after bits 1 and 2) of the cells whose primitives changed, bit 8: of every cell.
__global__ void __launch_bounds__(HFV_BLOCK) k_apply(Batch B, int mode) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  if (!B.eng[s].stepping) return;
  // fused step: a chunk with no active cell whose cells have no active neighbour (no chunk on its list of
  // neighbour chunks holds an active cell) has nothing to accumulate, apply or pack on a fixed mesh, and
  // nothing for the limiter sweeps (which read the decision from cflag)
  if (mode == 7) {
    bool work = B.P.act[J.chunk] > 0 || B.sims[s].cfg.motion != MOT_FIXED;
    if (!work) {  // block-uniform
      const int nn = B.F.nchn[J.chunk];
      const bool w = nn == 0xFF || ((int)threadIdx.x < nn && B.P.act[B.F.nch[(uint64_t)J.chunk * HFV_NCH_MAX + threadIdx.x]] > 0);
      work = __syncthreads_or(w);
    }
    if (threadIdx.x == 0) B.P.cflag[J.chunk] = work ? 1 : 0;
    if (!work) return;
  }
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  if (i >= J.hi) return;
  const Cells& C = B.C;
  const Faces& F = B.F;
  const int64_t n = C.n;
  const Cfg& cfg = B.sims[s].cfg;
  const bool f32 = cfg.precision == HFV_FP32_FLUX;
  const bool a = C.act[i];
  Events ev;
  bool have_prim = false;
  St<double> prim;
  if (mode & 1) {
    St<double> acc = vacuum<double>();
    bool have = false;
    double msv = 0., maor = 0.;
    if (a) {
      acc = ld5(C.facc, n, i);
      have = true;
      msv = C.msv[i];
      maor = C.maor[i];
    }
    const int32_t k0 = F.off[i], k1 = F.off[i + 1];
    // a face's record: its five fluxes, v_max and a_over_r
    auto load_rec = [&](uint32_t f, St<double>& r, double& vmx, double& aor) {
      if (f32) {
        const float4* r4 = (const float4*)((const float*)F.rec + 8 * (uint64_t)f);
        const float4 p = r4[0], q = r4[1];
        r = mkst((double)p.x, mk3((double)p.y, (double)p.z, (double)p.w), (double)q.x);
        vmx = q.y;
        aor = q.z;
      } else {
        const double2* r2 = (const double2*)(F.rec + 8 * (uint64_t)f);
        const double2 r01 = r2[0], r23 = r2[1], r45 = r2[2], r67 = r2[3];
        r = mkst(r01.x, mk3(r01.y, r23.x, r23.y), r45.x);
        vmx = r45.y;
        aor = r67.x;
      }
    };
    // Entries in chunks: a chunk's CSR values, then its faces' flags, then the records of its current
    // faces, all in flight together before the chunk's entries are applied in CSR order.
    for (int32_t kb = k0; kb < k1; kb += HFV_APPLY_CHUNK) {
      int32_t cv[HFV_APPLY_CHUNK];
#pragma unroll
      for (int u = 0; u < HFV_APPLY_CHUNK; ++u) cv[u] = kb + u < k1 ? F.conn[kb + u] : 0;
      uint32_t cf[HFV_APPLY_CHUNK];
#pragma unroll
      for (int u = 0; u < HFV_APPLY_CHUNK; ++u) cf[u] = kb + u < k1 ? (uint32_t)F.cur[(uint32_t)(cv[u] >> 1)] : 0u;
      St<double> rc[HFV_APPLY_CHUNK];
      double vm[HFV_APPLY_CHUNK], ao[HFV_APPLY_CHUNK];
#pragma unroll
      for (int u = 0; u < HFV_APPLY_CHUNK; ++u) {
        rc[u] = vacuum<double>();
        vm[u] = 0.;
        ao[u] = 0.;
        if (cf[u]) load_rec((uint32_t)(cv[u] >> 1), rc[u], vm[u], ao[u]);
      }
#pragma unroll
      for (int u = 0; u < HFV_APPLY_CHUNK; ++u) {
        if (!cf[u]) continue;
        if (!have) {
          acc = ld5(C.facc, n, i);
          have = true;
        }
        if ((cv[u] & 1) == 0)
          acc = acc - rc[u];
        else
          acc = acc + rc[u];
        if (a) {
          msv = hmax(vm[u], msv);
          maor = hmax(maor, ao[u]);
        }
      }
    }
    if (have) {
      if (a) {
        St<double> cc = ld5(C.cons, n, i) + acc;  // apply_flux (part.rs:178-193)
#pragma unroll
        for (int q = 0; q < 5; ++q)
          if (!hfinite(cc.c[q])) ev.fail(E_NONFINITE);
        if (cc.c[0] < 0.) {
          ev.flux_reset++;
          cc = vacuum<double>();
        }
        if (cc.c[4] < 0.) {
          ev.flux_reset++;
          cc = vacuum<double>();
        }
        acc = vacuum<double>();
        // kick2 -> grav_kick with a_grav = 0 (part.rs:303-312): momentum += (0 * mass)
        cc.c[1] += 0.0;
        cc.c[2] += 0.0;
        cc.c[3] += 0.0;
        st5(C.cons, n, i, cc);
        C.msv[i] = msv;
        C.maor[i] = maor;
        if (mode & 2) {
          prim = c2p(cc, C.vol[i], cfg.gd, ev);
          st5(C.prim, n, i, prim);
          have_prim = true;
        }
      }
      st5(C.facc, n, i, acc);
    }
  } else if ((mode & 2) && a) {
    prim = c2p(ld5(C.cons, n, i), C.vol[i], cfg.gd, ev);
    st5(C.prim, n, i, prim);
    have_prim = true;
  }
  // the pc pack changes where prim changed (active cells), or everywhere on moving meshes / on request
  if ((mode & 4) && (have_prim || (mode & 8) || cfg.motion != MOT_FIXED))
    store_pc(C.pc, i, have_prim ? prim : ld5(C.prim, n, i), ld3(C.cen, n, i));
  flush_events(ev, B.cnt, B.eng, s);
}

// The pc pack of cells [lo, hi) from the SoA fields (after the host changed the state or mesh).
__global__ void __launch_bounds__(HFV_BLOCK) k_pack_pc(Batch B, int64_t lo, int64_t hi) {
  const int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  store_pc(B.C.pc, i, ld5(B.C.prim, B.C.n, i), ld3(B.C.cen, B.C.n, i));
}

// The time-extrapolation rates of every loaded simulation's cells, and the flux records of the staged
// ones, from the SoA fields (after the host changed the state or the mesh, or a phase call changed
// fields on its own).
__global__ void __launch_bounds__(HFV_BLOCK) k_pack_fr(Batch B) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  if (i >= J.hi) return;
  const Cells& C = B.C;
  const int64_t n = C.n;
  const Cfg& cfg = B.sims[J.sim].cfg;
  const Gas<double>& gd = cfg.gd;
  const St<double> p = ld5(C.prim, n, i);
  const Gr<double> g = ld_grad(C, i, cfg.precision == HFV_FP32_FLUX);
  const St<double> rate = fr_rate(p, g, ld3(C.vrel, n, i), gd);
  st_tex(C, i, cfg.precision == HFV_FP32_FLUX, rate);
  if (B.sims[J.sim].staged)
    store_fr(C.fr, i, cfg.precision == HFV_FP32_FLUX, p, rate, fr_cs(p, gd), g, C.dt[i], ld3(C.gc, n, i));
}

// FP32-flux contexts: sum the unsummed predict time of the cells of simulation s (every loaded
// simulation for s < 0) into their extrapolations, with the rate of the current primitives, gradients
// and v_rel, before anything changes those (host writes, trait-level phases); flags the primitives for
// predict's check_physical.
__global__ void __launch_bounds__(HFV_BLOCK) k_settle(Batch B, int s) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0 || (s >= 0 && J.sim != s)) return;
  if (B.sims[J.sim].cfg.precision != HFV_FP32_FLUX) return;
  SimEngine& E = B.eng[J.sim];
  if (threadIdx.x == 0) E.chk = 1;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  if (i >= J.hi) return;
  const Cells& C = B.C;
  const double rho = C.prim[i];
  const double dt = (double)(E.ti_pred - C.tiex[i]) * E.time_base;
  st_ex(C, i, true, ld_ex(C, i, true) + (rho > 0. ? (-dt) * ld_tex(C, i, true) : vacuum<double>()));
  C.tiex[i] = E.ti_pred;
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

// KeplerianPotential::get_timestep (gravity.rs:120-123) with circular_velocity (gravity.rs:138-148):
// eta = 0.5 times the particle's 2D radius over sqrt(1 / r), r softened below the softening length.
__device__ __forceinline__ double kepler_timestep(const Cfg& cfg, V3<double> loc, double V) {
  const V3<double> dx = loc - mk3(cfg.grav_pos[0], cfg.grav_pos[1], cfg.grav_pos[2]);
  const double r = len(dx);
  const double r_soft = r < cfg.grav_soft ? r + cfg.grav_soft : r;
  return 0.5 * sqrt(0.318309886183790671537767526745028724 * V) / sqrt(1. / r_soft);
}

// DMat3::inverse (glam 0.27): cofactors by cross products, det = z . (x cross y), x 1/det, transposed;
// returned as the three columns of the inverse.
__device__ __forceinline__ M3 inverse_cols(const M3& m) {
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
  return inv;
}

// Space::get_boundary_part (space.rs:213-229): the mirrored neighbour of a wall face.
__device__ __forceinline__ void boundary_neighbour(const Faces& F, const Cfg& cfg, int64_t f, const V3<double>& cen,
                                                   const St<double>& prim, V3<double>& oc, St<double>& op) {
  const int64_t nf = F.nf;
  V3<double> fc = ld3(F.fcen, nf, f), nrm = ld3(F.nrm, nf, f);
  oc = cen + ((2. * dot(fc - cen, nrm)) * nrm);
  op = prim;
  if (cfg.boundary == BND_REFLECTIVE) op = reflect(prim, nrm);
  else if (cfg.boundary == BND_VACUUM) op = vacuum<double>();
}

// ------------------------------------------------------------------ k_grad
// gradient_estimate / gradient_limit / gradient_apply (hydrodynamics.rs:67-143, gradients.rs:8-103)
// fused with compute_timesteps (euler_eqns_fvs.rs:50-61, part.rs:42-84, 328-372) and timesteps_apply
// (hydrodynamics.rs:145-158, timeline.rs:103-139) for an active cell.  The limiter's neighbour min/max
// is gathered in the estimate sweep (same values, same order); the second sweep only extrapolates.
// STATIC (fixed 3D mesh): the per-entry weight w, offset ds and face offset dxf, and the inverse of the
// gradient matrix, are the values k_geom_cells computed once with the same expressions and order.  (On
// 1D/2D meshes recomputing them reads fewer bytes than the precomputed lines, so those stay dynamic.)
template <bool STATIC>
__device__ __forceinline__ void grad_cell(const Batch& B, const Cfg& cfg, const SimEngine& E, int64_t i, int do_grad,
                                          int do_ts, bool pack_fr, double* frv, bool v_zero, const double4* gst,
                                          uint64_t* dti_out, Events& ev) {
  const Cells& C = B.C;
  const Faces& F = B.F;
  const int64_t n = C.n, nf = F.nf;
  const St<double> prim = ld5(C.prim, n, i);
  const V3<double> cen = ld3(C.cen, n, i);
  Gr<double> gr;
  V3<double> gcen;
  const int32_t k0 = F.off[i], k1 = F.off[i + 1];
  if (do_grad && cfg.runner == 0) {
    // ---- estimate (GradientData) + neighbour extrema (LimiterData)
    Gr<double> g;
#pragma unroll
    for (int q = 0; q < 5; ++q) g.g[q] = mk3(0., 0., 0.);
    // the gradient matrix is symmetric bit for bit (w * (ds.a * ds.b) == w * (ds.b * ds.a), summed in the
    // same order), so its six distinct entries are accumulated
    double mxx = 0., mxy = 0., mxz = 0., myy = cfg.dim == 1 ? 1. : 0., myz = 0., mzz = cfg.dim <= 2 ? 1. : 0.;
    St<double> lmin = prim, lmax = prim;
    // one entry's terms: gradient sums (one component at a time: the scheduler would otherwise form every
    // product first) and the neighbour extrema
    auto accumulate = [&](const St<double>& op, const V3<double>& ds, double w) {
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        g.g[q] = g.g[q] + ((w * (op.c[q] - prim.c[q])) * ds);
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        lmin.c[q] = hmin(lmin.c[q], op.c[q]);
        lmax.c[q] = hmax(lmax.c[q], op.c[q]);
      }
      __builtin_amdgcn_sched_barrier(0);
    };
    if (STATIC) {
      // entries in chunks of HFV_GRAD_CHUNK: a chunk's neighbour references (LDS stage slot when staged,
      // else its cell; -1 at a wall) and geometry lines in flight together, then its entries in order
      for (int32_t kb = k0; kb < k1; kb += HFV_GRAD_CHUNK) {
        int32_t rr[HFV_GRAD_CHUNK];
        double4 ag[HFV_GRAD_CHUNK];
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          const int32_t k = kb + u;
          if (k < k1) {
            if (gst) {
              const uint32_t li = F.nloc[k];
              rr[u] = li == 0xFFFFu ? -1 : (int32_t)li;
            } else {
              rr[u] = F.nbr[k];
            }
            ag[u] = ((const double4*)(F.eg + 8 * (uint64_t)(uint32_t)k))[0];
          } else {
            rr[u] = -1;
            ag[u] = make_double4(0., 0., 0., 0.);
          }
        }
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          const int32_t k = kb + u;
          if (k >= k1) break;
          const int32_t other = rr[u];
          const double4 a = ag[u];
          const double w = a.x;
          const V3<double> ds = mk3(a.y, a.z, a.w);
          St<double> op;
          if (other >= 0) {
            V3<double> oc;
            if (gst) load_pc((const double*)(gst + 2 * (uint32_t)other), 0, op, oc);
            else load_pc(C.pc, (uint32_t)other, op, oc);
          } else {
            op = prim;  // Space::get_boundary_part (space.rs:213-229)
            if (cfg.boundary == BND_REFLECTIVE) op = reflect(prim, ld3(F.nrm, nf, (uint32_t)(F.conn[k] >> 1)));
            else if (cfg.boundary == BND_VACUUM) op = vacuum<double>();
          }
          accumulate(op, ds, w);
        }
      }
    } else {
      // Entries in chunks of HFV_GRAD_CHUNK: a chunk's CSR values and neighbour references, then its faces'
      // shift codes and areas, are all in flight together before its entries are accumulated in order, so
      // a cell waits for about two dependent loads per chunk instead of two per entry.
      auto ref_of = [&](int32_t k) -> int32_t {
        if (gst) {
          const uint32_t li = F.nloc[k];
          return li == 0xFFFFu ? -1 : (int32_t)li;
        }
        return F.nbr[k];
      };
      for (int32_t kb = k0; kb < k1; kb += HFV_GRAD_CHUNK) {
        int32_t ee[HFV_GRAD_CHUNK], rr[HFV_GRAD_CHUNK];
        uint32_t cc[HFV_GRAD_CHUNK];
        double aa[HFV_GRAD_CHUNK];
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          const bool in = kb + u < k1;
          ee[u] = in ? F.conn[kb + u] : 0;
          rr[u] = in ? ref_of(kb + u) : -1;
        }
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          const bool in = kb + u < k1;
          const uint32_t f = (uint32_t)(ee[u] >> 1);
          cc[u] = in ? (uint32_t)F.shc[f] : 0u;
          aa[u] = in ? ldo(F.area, f) : 0.;
        }
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          if (kb + u >= k1) break;
          const int32_t e = ee[u], other = rr[u];
          const double area = aa[u];
          St<double> op = prim;
          V3<double> oc = cen;
          if (other >= 0) {
            if (gst) load_pc((const double*)(gst + 2 * other), 0, op, oc);
            else load_pc(C.pc, (uint32_t)other, op, oc);
          }
          const uint32_t f = (uint32_t)(e >> 1);
          const V3<double> sh = decode_shift((uint8_t)cc[u], cfg.box);
          const V3<double> shift = (e & 1) == 0 ? sh : -sh;
          if (other < 0) boundary_neighbour(F, cfg, f, cen, prim, oc, op);
          const V3<double> ds = (oc + shift) - cen;
          const double w = area / len2(ds);
          __builtin_amdgcn_sched_barrier(0);
          mxx = mxx + w * (ds.x * ds.x);
          mxy = mxy + w * (ds.x * ds.y);
          mxz = mxz + w * (ds.x * ds.z);
          myy = myy + w * (ds.y * ds.y);
          myz = myz + w * (ds.y * ds.z);
          mzz = mzz + w * (ds.z * ds.z);
          __builtin_amdgcn_sched_barrier(0);
          accumulate(op, ds, w);
        }
      }
    }
    HFV_STAGE();
    // DMat3::inverse (glam 0.27): cofactors by cross products, det = z . (x cross y), transpose
    {
      M3 inv;
      if (STATIC) {
        inv.x = mk3(ldo(C.minv, (uint32_t)i), ldo(C.minv + n, (uint32_t)i), ldo(C.minv + 2 * n, (uint32_t)i));
        inv.y = mk3(ldo(C.minv + 3 * n, (uint32_t)i), ldo(C.minv + 4 * n, (uint32_t)i), ldo(C.minv + 5 * n, (uint32_t)i));
        inv.z = mk3(ldo(C.minv + 6 * n, (uint32_t)i), ldo(C.minv + 7 * n, (uint32_t)i), ldo(C.minv + 8 * n, (uint32_t)i));
      } else {
        M3 m;
        m.x = mk3(mxx, mxy, mxz);
        m.y = mk3(mxy, myy, myz);
        m.z = mk3(mxz, myz, mzz);
        inv = inverse_cols(m);
      }
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
    }
    HFV_STAGE();
    // ---- limit: extrema of the extrapolations to the faces
    St<double> emin, emax;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      emin.c[q] = INFINITY;
      emax.c[q] = -INFINITY;
    }
    auto extrema = [&](const V3<double>& dxf) {
      const St<double> ext = gdot(g, dxf);
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        emin.c[q] = hmin(emin.c[q], ext.c[q]);
        emax.c[q] = hmax(emax.c[q], ext.c[q]);
      }
    };
    if (STATIC) {
      for (int32_t kb = k0; kb < k1; kb += HFV_GRAD_CHUNK) {  // chunks of lines in flight together
        double4 bg[HFV_GRAD_CHUNK];
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u)
          bg[u] = kb + u < k1 ? ((const double4*)(F.eg + 8 * (uint64_t)(uint32_t)(kb + u)))[1] : make_double4(0., 0., 0., 0.);
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          if (kb + u >= k1) break;
          extrema(mk3(bg[u].x, bg[u].y, bg[u].z));
        }
      }
    } else {
      // chunked as the estimate sweep: a chunk's CSR values, then its faces' shift codes and centroids
      for (int32_t kb = k0; kb < k1; kb += HFV_GRAD_CHUNK) {
        int32_t ee[HFV_GRAD_CHUNK];
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) ee[u] = kb + u < k1 ? F.conn[kb + u] : 0;
        uint32_t cc[HFV_GRAD_CHUNK];
        V3<double> fcs[HFV_GRAD_CHUNK];
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          const bool in = kb + u < k1;
          const uint32_t f = (uint32_t)(ee[u] >> 1);
          cc[u] = in ? (uint32_t)F.shc[f] : 0u;
          fcs[u] = in ? ld3(F.fcen, nf, f) : cen;
        }
#pragma unroll
        for (int u = 0; u < HFV_GRAD_CHUNK; ++u) {
          if (kb + u >= k1) break;
          const V3<double> sh = decode_shift((uint8_t)cc[u], cfg.box);
          const V3<double> shift = (ee[u] & 1) == 0 ? sh : -sh;
          extrema((fcs[u] - cen) - shift);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      double d_max = lmax.c[q] - prim.c[q], d_min = lmin.c[q] - prim.c[q];
      double alpha = (emax.c[q] == 0. || emin.c[q] == 0.) ? 1. : hmin(1.0, hmin(d_max / emax.c[q], d_min / emin.c[q]));
      g.g[q] = alpha * g.g[q];
    }
    gr = round_grad(g, cfg.precision == HFV_FP32_FLUX);
    gcen = cen;
    st_grad(C, i, cfg.precision == HFV_FP32_FLUX, gr);
    st3(C.gc, n, i, gcen);
  } else {
    gr = ld_grad(C, i, cfg.precision == HFV_FP32_FLUX);
    gcen = ld3(C.gc, n, i);
  }
  if (!do_ts) return;
  HFV_STAGE();
  // ---- Particle::timestep (part.rs:42-84)
  const St<double> c = ld5(C.cons, n, i);
  double dt_req;
  V3<double> vrel_now;
  if (c.c[0] == 0.) {
    if (!v_zero) st3(C.v, n, i, mk3(0., 0., 0.));
    dt_req = INFINITY;
    vrel_now = ld3(C.vrel, n, i);
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
    // static geometry: fixed motion with every velocity +0 already, the value written here
    if (!v_zero) st3(C.v, n, i, v);
    st3(C.vrel, n, i, v_rel);
    vrel_now = v_rel;
    C.maor[i] = 0.;
    double v_max = hmax(C.msv[i], len(v_rel) + cs);
    C.msv[i] = 0.;
    dt_req = v_max > 0. ? cfg.cfl * cell_radius(V, cfg) / v_max : INFINITY;
  }
  // the minimum with the gravity solver's timestep (optimal_order.rs:47-50, godunov.rs:41-44); only the
  // Keplerian potential's is finite while a_grav = 0 (gravity.rs:45-54, 90-95)
  if (do_ts == 1 && cfg.gravity == HFV_GRAV_KEPLERIAN) dt_req = fmin(dt_req, kepler_timestep(cfg, ld3(C.loc, n, i), C.vol[i]));
  if (do_ts == 2) {  // FiniteVolumeSolver::compute_timesteps alone: the requested dt, not yet applied
    C.dtq[i] = dt_req;
    st_tex(C, i, cfg.precision == HFV_FP32_FLUX, fr_rate(prim, gr, vrel_now, cfg.gd));
    return;
  }
  // ---- timesteps_apply
  int8_t bin = C.bin[i];
  uint64_t dti = make_integer_timestep(dt_req, bin, HFV_NUM_TIME_BINS, E.ti_current, E.time_base_inv);
  const double dt_new = (double)dti * E.time_base;
  C.dt[i] = dt_new;
  C.bin[i] = get_time_bin(dti);
  *dti_out = dti;
  // the time-extrapolation rate and the flux record for the next step (prim, gradients, v_rel and dt
  // are final here)
  const St<double> rate_now = fr_rate(prim, gr, vrel_now, cfg.gd);
  st_tex(C, i, cfg.precision == HFV_FP32_FLUX, rate_now);
  if (pack_fr)
    fill_fr_any(frv, cfg.precision == HFV_FP32_FLUX, prim, rate_now, fr_cs(prim, cfg.gd), gr, dt_new, gcen);
}

template <bool STATIC>
__global__ void __launch_bounds__(HFV_BLOCK, HFV_GRAD_WAVES) k_grad(Batch B, int do_grad, int do_ts) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimEngine& E = B.eng[s];
  if (!E.stepping || B.sims[s].static_grad != (STATIC ? 1 : 0)) return;
  if (do_grad == 1 && do_ts == 1 && !B.P.act[J.chunk]) {  // fused step, no active cell in the chunk
    if (threadIdx.x == 0) B.P.ti[J.chunk] = ~0ull;
    return;
  }
  const Cfg& cfg = B.sims[s].cfg;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  const bool a = i < J.hi && B.C.act[i];
  uint64_t my = ~0ull;
  Events ev;
  const bool pack = do_ts == 1 && B.sims[s].staged;  // block-uniform
  // LDS: the staged neighbour lines of the gradient sweeps, then (once every wave is done with them)
  // the per-wave transposes of the new flux records
  __shared__ double4 lds4[HFV_GSTAGE_MAX * 2];
  static_assert(HFV_GSTAGE_MAX * 8 >= HFV_BLOCK * 17, "record transposes fit the stage buffer");
  const int gn = (do_grad && B.sims[s].gstaged) ? (int)B.F.gcnt[J.chunk] : 0;
  bool staged = false;
  if (gn > 0) {
    staged = __syncthreads_count(a) > 0;  // block-uniform
    if (staged) {
      // the chunk's list, then every line load of a round, issued together (two rounds)
      constexpr int NL = (HFV_GSTAGE_MAX * 4 + HFV_BLOCK - 1) / HFV_BLOCK;
      constexpr int NR = (NL + 1) / 2;
      const int32_t* gl = B.F.gcell + (uint64_t)J.chunk * HFV_GSTAGE_MAX;
      double2* d2 = (double2*)lds4;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        int32_t cell[NR];
        double2 v[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int k = (int)threadIdx.x + (r * NR + j) * HFV_BLOCK;
          cell[j] = (k >> 2) < gn ? gl[k >> 2] : 0;
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int k = (int)threadIdx.x + (r * NR + j) * HFV_BLOCK;
          v[j] = (k >> 2) < gn ? ((const double2*)(B.C.pc + 8 * (uint64_t)(uint32_t)cell[j]))[k & 3]
                               : make_double2(0., 0.);
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
          const int k = (int)threadIdx.x + (r * NR + j) * HFV_BLOCK;
          if ((k >> 2) < gn) d2[k] = v[j];
        }
      }
      __syncthreads();
    }
  }
  double frv[HFV_FR_W];
  if (a)
    grad_cell<STATIC>(B, cfg, E, i, do_grad, do_ts, pack, frv, B.sims[s].static_geom != 0, staged ? lds4 : nullptr,
                      &my, ev);
  if (pack) {
    if (staged) __syncthreads();  // the staged lines are dead from here on
    double* wbuf = (double*)lds4 + (threadIdx.x >> 6) * (64 * 17);
    const uint64_t has = __ballot(a);
    if (has)
      store_fr_wave(B.C.fr, i - (int64_t)__lane_id(), has, frv, wbuf, cfg.precision == HFV_FP32_FLUX ? 1 : 2);
  }
  if (do_ts == 1) {
    uint64_t t = block_min_u64(my);
    if (threadIdx.x == 0) B.P.ti[J.chunk] = t;
  }
  flush_events(ev, B.cnt, B.eng, s);
}

// timesteps_apply (hydrodynamics.rs:145-158) of the requested timesteps in dtq (the caller's, e.g. the
// minimum of the hydro and gravity timesteps, optimal_order.rs:43-47), for the active cells.
__global__ void __launch_bounds__(HFV_BLOCK) k_ts_apply(Batch B) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const SimEngine& E = B.eng[J.sim];
  if (!E.stepping) return;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  uint64_t my = ~0ull;
  if (i < J.hi && B.C.act[i]) {
    const uint64_t dti = make_integer_timestep(B.C.dtq[i], B.C.bin[i], HFV_NUM_TIME_BINS, E.ti_current, E.time_base_inv);
    B.C.dt[i] = (double)dti * E.time_base;
    B.C.bin[i] = get_time_bin(dti);
    my = dti;
  }
  const uint64_t t = block_min_u64(my);
  if (threadIdx.x == 0) B.P.ti[J.chunk] = t;
}

// GravitySolver::compute_timesteps (gravity.rs:6-10) of cells [lo, lo + n) of a simulation, a_grav = 0.
__global__ void __launch_bounds__(HFV_BLOCK) k_grav_dt(Batch B, Cfg cfg, int64_t lo, int64_t n, double* out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = lo + k;
  out[k] = cfg.gravity == HFV_GRAV_KEPLERIAN ? kepler_timestep(cfg, ld3(B.C.loc, B.C.n, i), B.C.vol[i]) : INFINITY;
}

// GravitySolver::compute_accelerations (gravity.rs:6-10) of cells [lo, lo + n); out [3n] component-major.
// Self gravity (gravity.rs:23-44): particle k sums m_j (x_j - x_k) / r^3, r = |x_j - x_k| + softening, over
// j = k+1 .. n-1 in increasing order (the reference's order, so sums agree bit for bit).  The block's tiles
// of positions and masses go through LDS; blocks of lower k have more partners and start first.
__global__ void __launch_bounds__(HFV_BLOCK) k_grav_acc(Batch B, Cfg cfg, int64_t lo, int64_t n, double* out) {
  __shared__ double tx[HFV_BLOCK], ty[HFV_BLOCK], tz[HFV_BLOCK], tm[HFV_BLOCK];
  const int64_t k0 = (int64_t)blockIdx.x * HFV_BLOCK;
  const int64_t k = k0 + threadIdx.x;
  const int64_t nc = B.C.n;
  const double* L = B.C.loc;
  V3<double> a = mk3(0., 0., 0.);
  if (cfg.gravity == HFV_GRAV_SELF) {
    const V3<double> xi = k < n ? ld3(L, nc, lo + k) : mk3(0., 0., 0.);
    for (int64_t t0 = k0; t0 < n; t0 += HFV_BLOCK) {  // block-uniform trip count
      const int64_t j = t0 + threadIdx.x;
      __syncthreads();
      if (j < n) {
        tx[threadIdx.x] = L[lo + j];
        ty[threadIdx.x] = L[nc + lo + j];
        tz[threadIdx.x] = L[2 * nc + lo + j];
        tm[threadIdx.x] = B.C.cons[lo + j];
      }
      __syncthreads();
      const int cnt = (int)(n - t0 < HFV_BLOCK ? n - t0 : HFV_BLOCK);
      if (k < n) {
        for (int q = 0; q < cnt; ++q) {
          if (t0 + q <= k) continue;
          const V3<double> dir = mk3(tx[q] - xi.x, ty[q] - xi.y, tz[q] - xi.z);
          const double r = len(dir) + cfg.grav_soft;
          const double r3 = r * r * r;
          a = a + mk3(tm[q] * dir.x / r3, tm[q] * dir.y / r3, tm[q] * dir.z / r3);
        }
      }
    }
  } else if (k < n) {
    if (cfg.gravity == HFV_GRAV_CONSTANT) {
      a = mk3(cfg.grav_acc[0], cfg.grav_acc[1], cfg.grav_acc[2]);
    } else {  // KeplerianPotential::acceleration (gravity.rs:125-136)
      const V3<double> dx = ld3(L, nc, lo + k) - mk3(cfg.grav_pos[0], cfg.grav_pos[1], cfg.grav_pos[2]);
      double r2 = dot(dx, dx);
      const double r = sqrt(r2);
      if (r < cfg.grav_soft) r2 = r2 + cfg.grav_soft * cfg.grav_soft;
      const double d = r * r2;
      a = mk3(-dx.x / d, -dx.y / d, -dx.z / d);
    }
  }
  if (k < n) {
    out[k] = a.x;
    out[n + k] = a.y;
    out[2 * n + k] = a.z;
  }
}

// ------------------------------------------------------------------ static geometry
// For a fixed mesh the geometric terms of the gradient sweeps and of the face signal speed depend only
// on the mesh and the (unchanging) generator positions: they are evaluated once, with the expressions
// of the per-step code, for cells [lo, hi) / faces [flo, fhi) of simulation s.
__global__ void __launch_bounds__(HFV_BLOCK) k_geom_cells(Batch B, int s, int64_t lo, int64_t hi) {
  const int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const Cells& C = B.C;
  const Faces& F = B.F;
  const Cfg& cfg = B.sims[s].cfg;
  const int64_t n = C.n, nf = F.nf;
  const V3<double> cen = ld3(C.cen, n, i);
  const St<double> prim = ld5(C.prim, n, i);
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
  const int32_t k0 = F.off[i], k1 = F.off[i + 1];
  for (int32_t k = k0; k < k1; ++k) {
    const int32_t e = F.conn[k];
    const uint32_t f = (uint32_t)(e >> 1);
    const V3<double> sh = decode_shift(F.shc[f], cfg.box);
    const V3<double> shift = (e & 1) == 0 ? sh : -sh;
    const int32_t other = F.nbr[k];
    V3<double> oc;
    St<double> op;
    if (other >= 0) oc = ld3(C.cen, n, other);
    else boundary_neighbour(F, cfg, f, cen, prim, oc, op);
    const V3<double> ds = (oc + shift) - cen;
    const double w = ldo(F.area, f) / len2(ds);
    m.x = m.x + (w * (ds.x * ds));
    m.y = m.y + (w * (ds.y * ds));
    m.z = m.z + (w * (ds.z * ds));
    const V3<double> dxf = (ld3(F.fcen, nf, f) - cen) - shift;
    double4* eg = (double4*)(F.eg + 8 * (uint64_t)(uint32_t)k);
    eg[0] = make_double4(w, ds.x, ds.y, ds.z);
    eg[1] = make_double4(dxf.x, dxf.y, dxf.z, 0.);
  }
  const M3 inv = inverse_cols(m);
  const double v[9] = {inv.x.x, inv.x.y, inv.x.z, inv.y.x, inv.y.y, inv.y.z, inv.z.x, inv.z.y, inv.z.z};
#pragma unroll
  for (int q = 0; q < 9; ++q) C.minv[q * n + i] = v[q];
}
__global__ void __launch_bounds__(HFV_BLOCK) k_geom_faces(Batch B, int s, int64_t flo, int64_t fhi) {
  const int64_t f = flo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= fhi) return;
  const Cells& C = B.C;
  const Faces& F = B.F;
  const Cfg& cfg = B.sims[s].cfg;
  const int64_t n = C.n, nf = F.nf;
  const int32_t l = F.left[f], r = F.right[f];
  if (l < 0) return;
  V3<double> dx_centroid, dx;
  if (r >= 0) {
    const V3<double> sh = decode_shift(F.shc[f], cfg.box);
    dx_centroid = (ld3(C.cen, n, r) + sh) - ld3(C.cen, n, l);
    dx = (ld3(C.loc, n, r) + sh) - ld3(C.loc, n, l);
  } else {
    const V3<double> fc = ld3(F.fcen, nf, f), nrm = ld3(F.nrm, nf, f);
    const V3<double> loc = ld3(C.loc, n, l), cen = ld3(C.cen, n, l);
    const V3<double> loc_r = loc + ((2. * dot(fc - loc, nrm)) * nrm);
    const V3<double> cen_r = cen + ((2. * dot(fc - cen, nrm)) * nrm);
    dx_centroid = cen_r - cen;
    dx = loc_r - loc;
  }
  *(double4*)(F.fs + 4 * (uint64_t)f) = make_double4(dx_centroid.x, dx_centroid.y, dx_centroid.z, F.area[f] / len(dx));
}

// ------------------------------------------------------------------ limiter
// timestep_limiter first sweep (hydrodynamics.rs:165-179): wakeup from active neighbours' new bins;
// timestep_sync (hydrodynamics.rs:186-191) reduces the minimum bin per simulation instead.
// A chunk without active cells or active neighbours (k_apply's decision in cflag): its wakeups are the
// cells' own bins and k_limit changes nothing there, so the fused step (fused = 1) skips it in both.
__device__ __forceinline__ bool limiter_idle(const Batch& B, const BlockJob& J) {
  return !B.sims[J.sim].cfg.sync_all && !B.P.cflag[J.chunk];
}

__global__ void __launch_bounds__(HFV_BLOCK) k_wakeup(Batch B, int fused) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  if (!B.eng[s].stepping) return;
  if (fused && limiter_idle(B, J)) return;
  const Cells& C = B.C;
  const Faces& F = B.F;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  const bool valid = i < J.hi;
  if (B.sims[s].cfg.sync_all) {
    int m = block_min_i32(valid ? (int)C.bin[i] : 0x7fffffff);
    if (threadIdx.x == 0) atomicMin(&B.eng[s].min_bin, m);
    return;
  }
  if (!valid) return;
  int8_t w = C.bin[i];
  const int32_t k0 = F.off[i], k1 = F.off[i + 1];
  // entries in chunks of 8: a chunk's neighbours, then their active flags and bins, all in flight together
  // (the bin is read whether or not the neighbour is active: same result, no dependent load; a minimum,
  // so the order does not matter)
  for (int32_t kb = k0; kb < k1; kb += 8) {
    int32_t o[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) o[u] = kb + u < k1 ? F.nbr[kb + u] : -1;
    uint32_t ab[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) ab[u] = o[u] >= 0 ? ((uint32_t)C.act[o[u]] << 8) | (uint8_t)C.bin[o[u]] : 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (!(ab[u] >> 8)) continue;
      const int8_t cand = (int8_t)((int8_t)(ab[u] & 0xFF) + 2);
      if (cand < w) w = cand;
    }
  }
  C.wake[i] = w;
}

// Particle::timestep_limit (part.rs:263-297) + reset_extrapolations (hydrodynamics.rs:193-195) +
// kick1 (hydro_kick1 part.rs:87-104, grav_kick with a_grav = 0).
// canon: apply kick1's "momentum += 0 * mass" (a no-op after kick2 of the same step, which the fused
// step has already done; the trait-level phases keep it).
__global__ void __launch_bounds__(HFV_BLOCK) k_limit(Batch B, int canon) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimEngine& E = B.eng[s];
  if (!E.stepping) return;
  if (!canon && limiter_idle(B, J)) return;
  const int64_t i = (int64_t)J.lo + threadIdx.x;
  if (i >= J.hi) return;
  const Cells& C = B.C;
  const Cfg& cfg = B.sims[s].cfg;
  const int64_t n = C.n;
  int8_t wakeup = cfg.sync_all ? (int8_t)E.min_bin : C.wake[i];
  int8_t bin = C.bin[i];
  if (wakeup < bin) {
    double dt = C.dt[i];
    uint64_t ti_current = E.ti_current;
    double time_base = E.time_base;
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
    if (B.sims[s].staged) C.fr[HFV_FR_W * (uint64_t)i + HFV_FR_DT] = dt;
  }
  if (C.act[i]) {
    st_ex(C, i, cfg.precision == HFV_FP32_FLUX, vacuum<double>());
    if (cfg.precision == HFV_FP32_FLUX) C.tiex[i] = E.ti_pred;  // the sum restarts at zero now
    if (cfg.motion != MOT_FIXED) {
      double rho = C.prim[i];
      if (rho > 0.) {
        double kick_fac = 0.5 * C.dt[i] / rho;
        const Gr<double> g = ld_grad(C, i, cfg.precision == HFV_FP32_FLUX);
        const double dp[3] = {g.g[4].x, g.g[4].y, g.g[4].z};
        for (int a = 0; a < cfg.dim; ++a) C.v[a * n + i] -= kick_fac * dp[a];
      }
    }
    if (canon) {
      C.cons[1 * n + i] += 0.0;
      C.cons[2 * n + i] += 0.0;
      C.cons[3 * n + i] += 0.0;
    }
  }
}

// Particle::from_ic + first_init (part.rs:200-217) for cells [lo, hi) of simulation s; vel3 is
// component-major [3m].
__global__ void k_init_ic(Batch B, int s, int64_t lo, int64_t hi, const double* mass, const double* vel3,
                          const double* u_int) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t i = lo + j;
  if (i >= hi) return;
  const Cells& C = B.C;
  const int64_t n = C.n, m_ = hi - lo;
  const Cfg& cfg = B.sims[s].cfg;
  Events ev;
  double m = mass[j];
  V3<double> v = mk3(vel3[j], vel3[m_ + j], vel3[2 * m_ + j]);
  St<double> c = mkst(m, m * v, m * u_int[j] + 0.5 * m * len2(v));
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
  flush_events(ev, B.cnt, B.eng, s);
}

// Space::status sums (space.rs:1248-1280) over cells [lo, hi): one block, fixed-order tree reduction.
__global__ void __launch_bounds__(1024) k_conservation(Batch B, int64_t lo, int64_t hi, double* out5) {
  __shared__ double sh[1024];
  const int64_t n = B.C.n;
  for (int q = 0; q < 5; ++q) {
    double s = 0.;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += B.C.cons[q * n + i] + B.C.facc[q * n + i];
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
