// C-ABI of the MI355X per-face finite-volume path (include/hfv.h) and the host engine loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/hfv.h"
#include "hfv_kernels.hpp"

using namespace hfv;

namespace {

thread_local std::string g_last_error;

int set_err(int code, const char* what) {
  g_last_error = what ? what : "";
  return code;
}

#define HIP_OK(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      char _b[512];                                                                           \
      snprintf(_b, sizeof(_b), "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
               __LINE__);                                                                     \
      return set_err(HFV_ERR_HIP, _b);                                                        \
    }                                                                                         \
  } while (0)

constexpr int BLOCK = 256;
inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

template <class T>
Gas<T> make_gas(double gamma, int eos, double iso_u) {
  // AdiabaticIndex (gas_law.rs:1-54), evaluated in double exactly as the reference does
  double gi = 1. / gamma;
  double odgm1 = 1. / (gamma - 1.);
  double odgp1 = 1. / (gamma + 1.);
  Gas<T> g;
  g.gamma = (T)gamma;
  g.gamma_inv = (T)gi;
  g.odgm1 = (T)odgm1;
  g.odgp1 = (T)odgp1;
  g.gp1dg = (T)((gamma + 1.) * gi);
  g.gm1d2g = (T)(0.5 * (gamma - 1.) * gi);
  g.gm1dgp1 = (T)((gamma - 1.) * odgp1);
  g.tdgm1 = (T)(2. * odgm1);
  g.tdgp1 = (T)(2. * odgp1);
  g.gamma_tdgm1 = (T)(gamma * (2. * odgm1));
  g.half_gm1 = (T)(0.5 * (gamma - 1.));
  g.one_over_gamma = (T)(1. / gamma);
  g.iso_u = (T)iso_u;
  g.isothermal = eos == HFV_ISOTHERMAL;
  return g;
}

Cfg make_cfg(const hfv_params& p) {
  Cfg c;
  c.dim = p.dim;
  c.boundary = p.boundary;
  c.motion = p.motion;
  c.runner = p.runner;
  c.sync_all = p.sync_all;
  for (int a = 0; a < 3; ++a) c.box[a] = p.box[a];
  c.cfl = p.cfl;
  c.radius3_fac = 0.25 * 3. * 0.318309886183790671537767526745028724;
  c.solver.kind = p.solver;
  c.solver.threshold = p.airs_threshold;
  for (int a = 0; a < 3; ++a) c.solver.linv[a] = p.linadv_velocity[a];
  c.gd = make_gas<double>(p.gamma, p.eos, p.isothermal_u);
  c.gf = make_gas<float>(p.gamma, p.eos, p.isothermal_u);
  return c;
}

int check_params(const hfv_params* p) {
  if (!p) return set_err(HFV_ERR_INVALID, "params is NULL");
  if (p->dim < 1 || p->dim > 3) return set_err(HFV_ERR_INVALID, "dim must be 1, 2 or 3");
  if (p->boundary < 0 || p->boundary > 3) return set_err(HFV_ERR_INVALID, "unknown boundary");
  if (p->solver < 0 || p->solver > 6) return set_err(HFV_ERR_INVALID, "unknown riemann solver");
  if (p->motion < 0 || p->motion > 3) return set_err(HFV_ERR_INVALID, "unknown particle motion");
  if (p->runner < 0 || p->runner > 1) return set_err(HFV_ERR_INVALID, "unknown runner");
  if (p->precision < 0 || p->precision > 1) return set_err(HFV_ERR_INVALID, "unknown precision");
  if (!(p->gamma > 1.)) return set_err(HFV_ERR_INVALID, "gamma must be > 1");
  if (!(p->t_end > 0.)) return set_err(HFV_ERR_INVALID, "t_end must be > 0");
  return HFV_OK;
}

template <class T>
T* dalloc(int64_t count, std::vector<void*>& owned) {
  void* p = nullptr;
  if (count <= 0) count = 1;
  if (hipMalloc(&p, sizeof(T) * (size_t)count) != hipSuccess) return nullptr;
  owned.push_back(p);
  return (T*)p;
}

}  // namespace

enum { SP_DUMP = 0, SP_STEP = 3 };

struct hfv_ctx {
  hfv_params p;
  Cfg cfg;
  int device;
  hipStream_t stream;
  int64_t n = 0, nf = 0, nconn = 0, face_cap = 0, conn_cap = 0;
  Cells C{};
  Faces F{};
  std::vector<void*> cell_bufs, face_bufs;
  DevCounters* dcnt = nullptr;
  DevCounters* hcnt = nullptr;  // pinned
  // engine (src/engine.rs)
  hfv_timestep_info ts{};
  double t_current = 0.;
  int64_t step_count = 0, cell_updates = 0, face_fluxes = 0, last_active = 0, last_face_active = 0;
  std::deque<std::pair<uint64_t, int>> sync_points;
  uint64_t ti_snap = 0, ti_between_snaps = 0;
  hfv_timestep_info last_info{};
  int64_t soft[4] = {0, 0, 0, 0};
  int err = 0;
  // kernel timing (hfv_profile_enable)
  bool prof = false;
  hipEvent_t ev0[HFV_NUM_KERNEL_SLOTS]{}, ev1[HFV_NUM_KERNEL_SLOTS]{};
  bool pending[HFV_NUM_KERNEL_SLOTS]{};
  hfv_kernel_times times{};
};

namespace {

int free_list(std::vector<void*>& v) {
  for (void* p : v) (void)hipFree(p);
  v.clear();
  return HFV_OK;
}

int alloc_cells(hfv_ctx* c, int64_t n) {
  free_list(c->cell_bufs);
  auto& o = c->cell_bufs;
  Cells& C = c->C;
  C.n = n;
  C.prim = dalloc<double>(5 * n, o);
  C.grad = dalloc<double>(15 * n, o);
  C.gc = dalloc<double>(3 * n, o);
  C.extrap = dalloc<double>(5 * n, o);
  C.cons = dalloc<double>(5 * n, o);
  C.facc = dalloc<double>(5 * n, o);
  C.vol = dalloc<double>(n, o);
  C.loc = dalloc<double>(3 * n, o);
  C.cen = dalloc<double>(3 * n, o);
  C.v = dalloc<double>(3 * n, o);
  C.vrel = dalloc<double>(3 * n, o);
  C.maor = dalloc<double>(n, o);
  C.msv = dalloc<double>(n, o);
  C.dt = dalloc<double>(n, o);
  C.srad = dalloc<double>(n, o);
  C.bin = dalloc<int8_t>(n, o);
  C.wake = dalloc<int8_t>(n, o);
  C.act = dalloc<uint8_t>(n, o);
  c->F.off = dalloc<int32_t>(n + 1, o);
  for (void* p : o)
    if (!p) return set_err(HFV_ERR_HIP, "hipMalloc failed (cells)");
  // zero every per-cell field (Particle::default, src/part.rs:10)
  const size_t szd = sizeof(double) * (size_t)n;
  double* dbl[] = {C.prim, C.grad, C.gc, C.extrap, C.cons, C.facc, C.vol, C.loc, C.cen, C.v, C.vrel, C.maor, C.msv, C.dt, C.srad};
  int mult[] = {5, 15, 3, 5, 5, 5, 1, 3, 3, 3, 3, 1, 1, 1, 1};
  for (int k = 0; k < 15; ++k) HIP_OK(hipMemsetAsync(dbl[k], 0, szd * mult[k], c->stream));
  HIP_OK(hipMemsetAsync(C.bin, 0, (size_t)n, c->stream));
  HIP_OK(hipMemsetAsync(C.wake, 0, (size_t)n, c->stream));
  HIP_OK(hipMemsetAsync(C.act, 0, (size_t)n, c->stream));
  c->n = n;
  return HFV_OK;
}

int alloc_faces(hfv_ctx* c, int64_t nf, int64_t nconn) {
  if (nf <= c->face_cap && nconn <= c->conn_cap && !c->face_bufs.empty()) return HFV_OK;
  free_list(c->face_bufs);
  auto& o = c->face_bufs;
  int64_t cap = std::max<int64_t>(nf, 1);
  int64_t ccap = std::max<int64_t>(nconn, 1);
  Faces& F = c->F;
  F.area = dalloc<double>(cap, o);
  F.nrm = dalloc<double>(3 * cap, o);
  F.fcen = dalloc<double>(3 * cap, o);
  F.rec = dalloc<double>(7 * cap, o);
  F.left = dalloc<int32_t>(cap, o);
  F.right = dalloc<int32_t>(cap, o);
  F.shc = dalloc<uint8_t>(cap, o);
  F.conn = dalloc<int32_t>(ccap, o);
  for (void* p : o)
    if (!p) return set_err(HFV_ERR_HIP, "hipMalloc failed (faces)");
  HIP_OK(hipMemsetAsync(F.rec, 0, sizeof(double) * 7 * (size_t)cap, c->stream));
  c->face_cap = cap;
  c->conn_cap = ccap;
  return HFV_OK;
}

int sync_and_check(hfv_ctx* c) {
  HIP_OK(hipMemcpyAsync(c->hcnt, c->dcnt, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (c->hcnt->err) {
    c->err = c->hcnt->err;
    char b[128];
    snprintf(b, sizeof(b), "device raised hard error %d", c->hcnt->err);
    return set_err(c->hcnt->err, b);
  }
  return HFV_OK;
}

int8_t max_active_bin(uint64_t ti) { return get_max_active_bin(ti); }

enum { KS_PREDICT = 0, KS_MESH1D, KS_FLUX, KS_APPLY, KS_GRAD, KS_WAKEUP, KS_LIMIT, KS_DRIFT };

inline void prof_begin(hfv_ctx* c, int slot) {
  if (c->prof) (void)hipEventRecord(c->ev0[slot], c->stream);
}
inline void prof_end(hfv_ctx* c, int slot) {
  if (c->prof) {
    (void)hipEventRecord(c->ev1[slot], c->stream);
    c->pending[slot] = true;
  }
}
// Called after the stream was synchronized.
void prof_collect(hfv_ctx* c) {
  if (!c->prof) return;
  for (int k = 0; k < HFV_NUM_KERNEL_SLOTS; ++k)
    if (c->pending[k]) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->ev0[k], c->ev1[k]) == hipSuccess) {
        c->times.ms[k] += ms;
        c->times.launches[k] += 1;
      }
      c->pending[k] = false;
    }
}

int launch_flux(hfv_ctx* c) {
  if (c->nf == 0) return HFV_OK;
  prof_begin(c, KS_FLUX);
  if (c->p.precision == HFV_FP32_FLUX)
    hipLaunchKernelGGL(k_flux<float>, dim3(grid_for(c->nf)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt);
  else
    hipLaunchKernelGGL(k_flux<double>, dim3(grid_for(c->nf)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt);
  prof_end(c, KS_FLUX);
  HIP_OK(hipGetLastError());
  return HFV_OK;
}

int mesh_update_1d(hfv_ctx* c) {
  if (c->p.dim != 1) return set_err(HFV_ERR_UNSUPPORTED, "device mesh rebuild is 1D only; rebuild 2D/3D meshes on the host");
  int periodic = c->p.boundary == HFV_PERIODIC;
  // the kernels assume the 1D face layout of the tessellator: face k between cells k-1 and k
  if (c->nf != (periodic ? c->n : c->n + 1))
    return set_err(HFV_ERR_INVALID, "1D moving mesh needs the 1D tessellator's face layout");
  prof_begin(c, KS_MESH1D);
  hipLaunchKernelGGL(k_voronoi1d_faces, dim3(grid_for(c->nf)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, periodic, c->dcnt);
  hipLaunchKernelGGL(k_voronoi1d_cells, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, periodic);
  prof_end(c, KS_MESH1D);
  HIP_OK(hipGetLastError());
  return HFV_OK;
}

// Kernels of one runner step after predict/mesh: flux .. limiter + kick1.
int launch_step_tail(hfv_ctx* c, const hfv_timestep_info* info) {
  int e = launch_flux(c);
  if (e) return e;
  prof_begin(c, KS_APPLY);
  hipLaunchKernelGGL(k_apply_c2p, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt, 3);
  prof_end(c, KS_APPLY);
  prof_begin(c, KS_GRAD);
  hipLaunchKernelGGL(k_grad_timestep, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt,
                     info->ti_current, info->time_base, info->time_base_inv, 1, 1);
  prof_end(c, KS_GRAD);
  prof_begin(c, KS_WAKEUP);
  hipLaunchKernelGGL(k_wakeup, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->dcnt, c->p.sync_all);
  prof_end(c, KS_WAKEUP);
  prof_begin(c, KS_LIMIT);
  hipLaunchKernelGGL(k_limit_kick, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->cfg, c->dcnt,
                     c->p.sync_all, info->ti_current, info->time_base);
  prof_end(c, KS_LIMIT);
  HIP_OK(hipGetLastError());
  return HFV_OK;
}

int finish_step(hfv_ctx* c, const hfv_timestep_info* info, uint64_t* ti_next) {
  int e = sync_and_check(c);
  prof_collect(c);
  if (e) return e;
  uint64_t m = c->hcnt->ti_min;
  if (c->hcnt->n_active == 0 || m == ~0ull) return set_err(HFV_ERR_INVALID, "no active cell in this step");
  c->last_active = (int64_t)c->hcnt->n_active;
  c->last_face_active = (int64_t)c->hcnt->n_face_active;
  if (ti_next) *ti_next = info->ti_current + m;
  return HFV_OK;
}

int require_mesh(hfv_ctx* c) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  if (c->n == 0 || c->face_bufs.empty()) return set_err(HFV_ERR_INVALID, "mesh not set");
  HIP_OK(hipSetDevice(c->device));
  return HFV_OK;
}

}  // namespace

// ============================================================== C-ABI
extern "C" {

const char* hfv_last_error(void) { return g_last_error.c_str(); }

int hfv_device_count(int* n) {
  if (!n) return set_err(HFV_ERR_INVALID, "n is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *n = 0;
    return set_err(HFV_ERR_NO_DEVICE, hipGetErrorString(e));
  }
  *n = c;
  return HFV_OK;
}

int hfv_create(const hfv_params* params, hfv_ctx** out) {
  int e = check_params(params);
  if (e) return e;
  if (!out) return set_err(HFV_ERR_INVALID, "out is NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(HFV_ERR_NO_DEVICE, "no HIP device");
  hfv_ctx* c = new hfv_ctx();
  c->p = *params;
  c->cfg = make_cfg(*params);
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "hipGetDevice failed");
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "hipStreamCreate failed");
  }
  if (hipMalloc((void**)&c->dcnt, sizeof(DevCounters)) != hipSuccess ||
      hipHostMalloc((void**)&c->hcnt, sizeof(DevCounters), hipHostMallocDefault) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "counter allocation failed");
  }
  hipLaunchKernelGGL(k_clear_all_counters, dim3(1), dim3(1), 0, c->stream, c->dcnt);
  // Engine::new (engine.rs:104-149)
  c->ts.time_base = params->t_end / (double)HFV_MAX_NR_TIMESTEPS;
  c->ts.time_base_inv = 1. / c->ts.time_base;
  c->sync_points.push_back({0, SP_STEP});
  c->ti_snap = 0;
  c->ti_between_snaps = params->dt_snap > 0. ? sat_u64(params->dt_snap * c->ts.time_base_inv) : HFV_MAX_NR_TIMESTEPS;
  if (c->ti_between_snaps == 0) c->ti_between_snaps = HFV_MAX_NR_TIMESTEPS;
  *out = c;
  return HFV_OK;
}

int hfv_destroy(hfv_ctx* c) {
  if (!c) return HFV_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_list(c->cell_bufs);
  free_list(c->face_bufs);
  for (int k = 0; k < HFV_NUM_KERNEL_SLOTS; ++k) {
    if (c->ev0[k]) (void)hipEventDestroy(c->ev0[k]);
    if (c->ev1[k]) (void)hipEventDestroy(c->ev1[k]);
  }
  if (c->dcnt) (void)hipFree(c->dcnt);
  if (c->hcnt) (void)hipHostFree(c->hcnt);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return HFV_OK;
}

int hfv_set_mesh(hfv_ctx* c, const hfv_mesh_arrays* m, const uint8_t* update_mask) {
  if (!c || !m) return set_err(HFV_ERR_INVALID, "NULL argument");
  HIP_OK(hipSetDevice(c->device));
  int64_t n = m->n_cells, nf = m->n_faces;
  if (n <= 0 || nf < 0) return set_err(HFV_ERR_INVALID, "empty mesh");
  if (n >= (1ll << 30) || nf >= (1ll << 29)) return set_err(HFV_ERR_INVALID, "mesh too large for 32-bit indices");
  if (c->n != 0 && c->n != n) return set_err(HFV_ERR_INVALID, "cell count differs from the existing state");
  bool first = c->n == 0;
  if (first) {
    int e = alloc_cells(c, n);
    if (e) return e;
  }
  // faces -> SoA, periodic shift -> code
  std::vector<double> area(nf), nrm(3 * nf), fcen(3 * nf);
  std::vector<int32_t> left(nf), right(nf);
  std::vector<uint8_t> shc(nf);
  for (int64_t f = 0; f < nf; ++f) {
    area[f] = m->area[f];
    left[f] = (int32_t)m->left[f];
    right[f] = (int32_t)m->right[f];
    if (m->left[f] < 0 || m->left[f] >= n || m->right[f] < -1 || m->right[f] >= n)
      return set_err(HFV_ERR_INVALID, "face cell index out of range");
    uint8_t code = 0;
    for (int a = 0; a < 3; ++a) {
      nrm[a * nf + f] = m->normal[3 * f + a];
      fcen[a * nf + f] = m->centroid[3 * f + a];
      double s = m->shift ? m->shift[3 * f + a] : 0.;
      int cc;
      if (s == 0.) cc = 0;
      else if (s == c->p.box[a]) cc = 1;
      else if (s == -c->p.box[a]) cc = 2;
      else return set_err(HFV_ERR_INVALID, "face shift is not 0 or +-box");
      code |= (uint8_t)(cc << (2 * a));
    }
    shc[f] = code;
  }
  // CSR: repack in cell order, encode side
  std::vector<int32_t> off(n + 1);
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) {
    off[i] = (int32_t)tot;
    tot += m->count[i];
  }
  off[n] = (int32_t)tot;
  if (tot >= (1ll << 31)) return set_err(HFV_ERR_INVALID, "csr too large");
  std::vector<int32_t> conn(tot);
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t k = 0; k < m->count[i]; ++k) {
      int64_t idx = m->offset[i] + k;
      if (idx < 0 || idx >= m->n_conn) return set_err(HFV_ERR_INVALID, "csr offset out of range");
      int64_t f = m->conn[idx];
      if (f < 0 || f >= nf) return set_err(HFV_ERR_INVALID, "csr face index out of range");
      int side;
      if (m->left[f] == i) side = 0;
      else if (m->right[f] == i) side = 1;
      else return set_err(HFV_ERR_INVALID, "csr lists a face that does not touch the cell");
      conn[off[i] + k] = (int32_t)((f << 1) | side);
    }
  }
  int e = alloc_faces(c, nf, tot);
  if (e) return e;
  c->nf = nf;
  c->nconn = tot;
  c->F.nf = nf;
  Faces& F = c->F;
  hipStream_t s = c->stream;
  HIP_OK(hipMemcpyAsync(F.area, area.data(), sizeof(double) * nf, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.nrm, nrm.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.fcen, fcen.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.left, left.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.right, right.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.shc, shc.data(), nf, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.off, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(F.conn, conn.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, s));
  // Particle::update_geometry for the selected cells
  std::vector<double> vol(n), cen(3 * n), srad(n);
  std::vector<double> old_vol, old_cen, old_srad;
  bool partial = update_mask && !first;
  if (partial) {
    old_vol.resize(n);
    old_cen.resize(3 * n);
    old_srad.resize(n);
    HIP_OK(hipMemcpyAsync(old_vol.data(), c->C.vol, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(old_cen.data(), c->C.cen, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(old_srad.data(), c->C.srad, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  }
  for (int64_t i = 0; i < n; ++i) {
    bool upd = !partial || update_mask[i];
    vol[i] = upd ? m->volume[i] : old_vol[i];
    srad[i] = upd ? m->search_radius[i] : old_srad[i];
    for (int a = 0; a < 3; ++a) cen[a * n + i] = upd ? m->cell_centroid[3 * i + a] : old_cen[a * n + i];
  }
  HIP_OK(hipMemcpyAsync(c->C.vol, vol.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(c->C.cen, cen.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(c->C.srad, srad.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIP_OK(hipStreamSynchronize(s));
  return HFV_OK;
}

int hfv_init_from_ic(hfv_ctx* c, int64_t n, const double* loc, const double* mass, const double* vel,
                     const double* u_int) {
  int e = require_mesh(c);
  if (e) return e;
  if (n != c->n || !loc || !mass || !vel || !u_int) return set_err(HFV_ERR_INVALID, "bad IC arrays");
  std::vector<double> locs(3 * n), vels(3 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      locs[a * n + i] = loc[3 * i + a];
      vels[a * n + i] = vel[3 * i + a];
    }
  std::vector<void*> tmp;
  double* dm = dalloc<double>(n, tmp);
  double* dv = dalloc<double>(3 * n, tmp);
  double* du = dalloc<double>(n, tmp);
  if (!dm || !dv || !du) {
    free_list(tmp);
    return set_err(HFV_ERR_HIP, "hipMalloc failed");
  }
  hipStream_t s = c->stream;
  HIP_OK(hipMemcpyAsync(c->C.loc, locs.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(dm, mass, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(dv, vels.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(du, u_int, sizeof(double) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_init_ic, dim3(grid_for(n)), dim3(BLOCK), 0, s, c->C, c->cfg, dm, dv, du, c->dcnt);
  HIP_OK(hipGetLastError());
  e = sync_and_check(c);
  free_list(tmp);
  return e;
}

#define STATE_FIELDS(X)                                                                           \
  X(prim, C.prim, 5) X(grad, C.grad, 15) X(gc, C.gc, 3) X(extrap, C.extrap, 5) X(cons, C.cons, 5) \
  X(fluxes, C.facc, 5) X(volume, C.vol, 1) X(loc, C.loc, 3) X(centroid, C.cen, 3) X(v, C.v, 3)    \
  X(v_rel, C.vrel, 3) X(maor, C.maor, 1) X(msv, C.msv, 1) X(dt, C.dt, 1)

int hfv_set_state(hfv_ctx* c, const hfv_state* st) {
  int e = require_mesh(c);
  if (e) return e;
  if (!st) return set_err(HFV_ERR_INVALID, "state is NULL");
  const int64_t n = c->n;
  hipStream_t s = c->stream;
#define SETF(name, dst, w) \
  if (st->name) HIP_OK(hipMemcpyAsync(c->dst, st->name, sizeof(double) * (w) * n, hipMemcpyHostToDevice, s));
  STATE_FIELDS(SETF)
#undef SETF
  if (st->timebin) HIP_OK(hipMemcpyAsync(c->C.bin, st->timebin, (size_t)n, hipMemcpyHostToDevice, s));
  HIP_OK(hipStreamSynchronize(s));
  return HFV_OK;
}

int hfv_get_state(hfv_ctx* c, hfv_state* st) {
  int e = require_mesh(c);
  if (e) return e;
  if (!st) return set_err(HFV_ERR_INVALID, "state is NULL");
  const int64_t n = c->n;
  hipStream_t s = c->stream;
#define GETF(name, src, w) \
  if (st->name) HIP_OK(hipMemcpyAsync(st->name, c->src, sizeof(double) * (w) * n, hipMemcpyDeviceToHost, s));
  STATE_FIELDS(GETF)
#undef GETF
  if (st->timebin) HIP_OK(hipMemcpyAsync(st->timebin, c->C.bin, (size_t)n, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return HFV_OK;
}

int hfv_get_stats(hfv_ctx* c, hfv_stats* out) {
  if (!c || !out) return set_err(HFV_ERR_INVALID, "NULL argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipMemcpyAsync(c->hcnt, c->dcnt, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  out->check_physical_resets = (int64_t)c->hcnt->check_physical;
  out->apply_flux_resets = (int64_t)c->hcnt->flux_reset;
  out->gradient_fallbacks = (int64_t)c->hcnt->grad_fallback;
  out->brent_calls = (int64_t)c->hcnt->brent;
  out->error_code = c->hcnt->err;
  return HFV_OK;
}

int hfv_synchronize(hfv_ctx* c) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  return HFV_OK;
}

// ---- FiniteVolumeSolver trait surface
int hfv_predict(hfv_ctx* c, double dt) {
  int e = require_mesh(c);
  if (e) return e;
  int8_t mab = max_active_bin(c->last_info.ti_current);
  hipLaunchKernelGGL(k_mask_predict, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->cfg, dt, mab, 1, c->dcnt);
  HIP_OK(hipGetLastError());
  return sync_and_check(c);
}

int hfv_compute_fluxes(hfv_ctx* c, const hfv_timestep_info* info) {
  int e = require_mesh(c);
  if (e) return e;
  if (!info) return set_err(HFV_ERR_INVALID, "info is NULL");
  c->last_info = *info;
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, c->stream, c->dcnt);
  hipLaunchKernelGGL(k_mask_predict, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->cfg, 0.,
                     max_active_bin(info->ti_current), 0, c->dcnt);
  e = launch_flux(c);
  if (e) return e;
  return sync_and_check(c);
}

int hfv_get_fluxes(hfv_ctx* c, double* out7f) {
  int e = require_mesh(c);
  if (e) return e;
  if (!out7f) return set_err(HFV_ERR_INVALID, "out is NULL");
  const int64_t nf = c->nf;
  std::vector<double> soa(7 * nf);
  HIP_OK(hipMemcpyAsync(soa.data(), c->F.rec, sizeof(double) * 7 * nf, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int64_t f = 0; f < nf; ++f)
    for (int k = 0; k < 7; ++k) out7f[7 * f + k] = soa[k * nf + f];
  return HFV_OK;
}

int hfv_apply_fluxes(hfv_ctx* c) {
  int e = require_mesh(c);
  if (e) return e;
  hipLaunchKernelGGL(k_apply_c2p, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt, 1);
  HIP_OK(hipGetLastError());
  return sync_and_check(c);
}

int hfv_conserved_to_primitive(hfv_ctx* c) {
  int e = require_mesh(c);
  if (e) return e;
  hipLaunchKernelGGL(k_apply_c2p, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt, 2);
  HIP_OK(hipGetLastError());
  return sync_and_check(c);
}

int hfv_gradients(hfv_ctx* c) {
  int e = require_mesh(c);
  if (e) return e;
  const hfv_timestep_info& t = c->last_info;
  hipLaunchKernelGGL(k_grad_timestep, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt,
                     t.ti_current, t.time_base, t.time_base_inv, 1, 0);
  HIP_OK(hipGetLastError());
  return sync_and_check(c);
}

int hfv_timesteps(hfv_ctx* c, uint64_t* ti_next) {
  int e = require_mesh(c);
  if (e) return e;
  const hfv_timestep_info& t = c->last_info;
  hipLaunchKernelGGL(k_grad_timestep, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->cfg, c->dcnt,
                     t.ti_current, t.time_base, t.time_base_inv, 0, 1);
  hipLaunchKernelGGL(k_wakeup, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->F, c->dcnt, c->p.sync_all);
  hipLaunchKernelGGL(k_limit_kick, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->cfg, c->dcnt,
                     c->p.sync_all, t.ti_current, t.time_base);
  HIP_OK(hipGetLastError());
  return finish_step(c, &t, ti_next);
}

// ---- fused step
int hfv_step(hfv_ctx* c, const hfv_timestep_info* info, uint64_t* ti_next) {
  int e = require_mesh(c);
  if (e) return e;
  if (!info) return set_err(HFV_ERR_INVALID, "info is NULL");
  c->last_info = *info;
  double dt_pred = (double)(info->ti_current - info->ti_old) * info->time_base;
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, c->stream, c->dcnt);
  prof_begin(c, KS_PREDICT);
  hipLaunchKernelGGL(k_mask_predict, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->cfg, dt_pred,
                     max_active_bin(info->ti_current), 1, c->dcnt);
  prof_end(c, KS_PREDICT);
  if (c->p.motion != HFV_FIXED) {
    e = mesh_update_1d(c);
    if (e) return e;
  }
  e = launch_step_tail(c, info);
  if (e) return e;
  return finish_step(c, info, ti_next);
}

// ---- engine (src/engine.rs:152-268)
int hfv_run(hfv_ctx* c, int64_t max_steps, int64_t* steps_done) {
  int e = require_mesh(c);
  if (e) return e;
  int64_t done = 0;
  while (c->t_current < c->p.t_end && done < max_steps) {
    if (c->sync_points.empty()) return set_err(HFV_ERR_INVALID, "sync point queue empty");
    auto sp = c->sync_points.front();
    c->sync_points.pop_front();
    if (sp.first < c->ts.ti_current) return set_err(HFV_ERR_INVALID, "drift backwards in time");
    uint64_t dti = sp.first - c->ts.ti_current;
    if (dti > 0) {
      double dt = (double)dti * c->ts.time_base;
      if (c->p.motion != HFV_FIXED) {  // with fixed motion v == 0 and the drift is the identity
        prof_begin(c, KS_DRIFT);
        hipLaunchKernelGGL(k_drift, dim3(grid_for(c->n)), dim3(BLOCK), 0, c->stream, c->C, c->cfg, dt, c->dcnt);
        prof_end(c, KS_DRIFT);
        HIP_OK(hipGetLastError());
      }
      c->ts.ti_current += dti;
      c->t_current = (double)c->ts.ti_current * c->ts.time_base;
    }
    if (sp.second == SP_DUMP) continue;
    uint64_t ti_next = 0;
    e = hfv_step(c, &c->ts, &ti_next);
    if (e) return e;
    c->cell_updates += c->last_active;
    c->face_fluxes += c->last_face_active;
    while (c->ti_snap < ti_next) {
      c->sync_points.push_back({c->ti_snap, SP_DUMP});
      c->ti_snap += c->ti_between_snaps;
    }
    c->sync_points.push_back({ti_next, SP_STEP});
    c->ts.ti_old = c->ts.ti_next;
    c->ts.ti_next = ti_next;
    c->step_count++;
    done++;
  }
  if (steps_done) *steps_done = done;
  return HFV_OK;
}

int hfv_get_engine_info(hfv_ctx* c, hfv_engine_info* out) {
  if (!c || !out) return set_err(HFV_ERR_INVALID, "NULL argument");
  out->ts = c->ts;
  out->t_current = c->t_current;
  out->t_end = c->p.t_end;
  out->step_count = c->step_count;
  out->cell_updates = c->cell_updates;
  out->face_fluxes = c->face_fluxes;
  out->last_n_active = c->last_active;
  out->last_n_face_active = c->last_face_active;
  out->done = !(c->t_current < c->p.t_end);
  return HFV_OK;
}

int hfv_conservation(hfv_ctx* c, double* out5) {
  int e = require_mesh(c);
  if (e) return e;
  if (!out5) return set_err(HFV_ERR_INVALID, "out is NULL");
  std::vector<void*> tmp;
  double* d = dalloc<double>(5, tmp);
  if (!d) return set_err(HFV_ERR_HIP, "hipMalloc failed");
  hipLaunchKernelGGL(k_conservation, dim3(1), dim3(1024), 0, c->stream, c->C, d);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out5, d, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  free_list(tmp);
  return HFV_OK;
}

int hfv_profile_enable(hfv_ctx* c, int enable) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (enable) {
    for (int k = 0; k < HFV_NUM_KERNEL_SLOTS; ++k) {
      if (!c->ev0[k]) HIP_OK(hipEventCreate(&c->ev0[k]));
      if (!c->ev1[k]) HIP_OK(hipEventCreate(&c->ev1[k]));
      c->pending[k] = false;
    }
    c->times = hfv_kernel_times{};
  }
  c->prof = enable != 0;
  return HFV_OK;
}

int hfv_profile_read(hfv_ctx* c, hfv_kernel_times* out) {
  if (!c || !out) return set_err(HFV_ERR_INVALID, "NULL argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  prof_collect(c);
  *out = c->times;
  return HFV_OK;
}

int hfv_riemann_batch(const hfv_params* params, int64_t nb, const double* WL, const double* WR, const double* vface,
                      const double* nrm, double* Fo) {
  int e = check_params(params);
  if (e) return e;
  if (nb <= 0 || !WL || !WR || !vface || !nrm || !Fo) return set_err(HFV_ERR_INVALID, "bad batch arrays");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(HFV_ERR_NO_DEVICE, "no HIP device");
  Cfg cfg = make_cfg(*params);
  std::vector<void*> tmp;
  double* dl = dalloc<double>(5 * nb, tmp);
  double* dr = dalloc<double>(5 * nb, tmp);
  double* dv = dalloc<double>(3 * nb, tmp);
  double* dn = dalloc<double>(3 * nb, tmp);
  double* df = dalloc<double>(5 * nb, tmp);
  DevCounters* cnt = dalloc<DevCounters>(1, tmp);
  if (!dl || !dr || !dv || !dn || !df || !cnt) {
    free_list(tmp);
    return set_err(HFV_ERR_HIP, "hipMalloc failed");
  }
  DevCounters h{};
  HIP_OK(hipMemcpy(dl, WL, sizeof(double) * 5 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dr, WR, sizeof(double) * 5 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dv, vface, sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dn, nrm, sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(cnt, &h, sizeof(DevCounters), hipMemcpyHostToDevice));
  if (params->precision == HFV_FP32_FLUX)
    hipLaunchKernelGGL(k_riemann_batch<float>, dim3(grid_for(nb)), dim3(BLOCK), 0, 0, cfg, nb, dl, dr, dv, dn, df, cnt);
  else
    hipLaunchKernelGGL(k_riemann_batch<double>, dim3(grid_for(nb)), dim3(BLOCK), 0, 0, cfg, nb, dl, dr, dv, dn, df, cnt);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(Fo, df, sizeof(double) * 5 * nb, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&h, cnt, sizeof(DevCounters), hipMemcpyDeviceToHost));
  free_list(tmp);
  if (h.err) return set_err(h.err, "device raised a hard error in the Riemann solver");
  return HFV_OK;
}

}  // extern "C"
