// C-ABI of the MI355X per-face finite-volume path (include/hfv.h).
//
// A context is a batch of slots (hfv_create makes a batch of one).  Device-resident engines advance
// in HBM; the host only enqueues launch sequences (replayed from a captured HIP graph) and reads the
// engines back every few steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hfv.h"
#include "hfv_kernels.hpp"
#include "hfv_voronoi2d.hpp"
#include "hfv_voronoi3d.hpp"

using namespace hfv;

namespace {

thread_local std::string g_last_error;

int set_err(int code, const char* what) {
  g_last_error = what ? what : "";
  return code;
}

#define HIP_OK(expr)                                                                               \
  do {                                                                                             \
    hipError_t _e = (expr);                                                                        \
    if (_e != hipSuccess) {                                                                        \
      char _b[512];                                                                                \
      snprintf(_b, sizeof(_b), "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
      return set_err(HFV_ERR_HIP, _b);                                                             \
    }                                                                                              \
  } while (0)

#define TRY(expr)            \
  do {                       \
    int _r = (expr);         \
    if (_r != HFV_OK) return _r; \
  } while (0)

constexpr int BLOCK = HFV_BLOCK;
inline unsigned grid_full(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK); }
inline int64_t chunks_of(int64_t items) { return std::max<int64_t>(1, (items + BLOCK - 1) / BLOCK); }

template <class T>
Gas<T> make_gas(double gamma, int eos, double iso_u) {
  // AdiabaticIndex (gas_law.rs:1-54), evaluated in double with the reference's expressions
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
  std::memset(&c, 0, sizeof(c));
  c.dim = p.dim;
  c.boundary = p.boundary;
  c.motion = p.motion;
  c.runner = p.runner;
  c.sync_all = p.sync_all;
  c.precision = p.precision;
  for (int a = 0; a < 3; ++a) c.box[a] = p.box[a];
  c.cfl = p.cfl;
  c.radius3_fac = 0.25 * 3. * 0.318309886183790671537767526745028724;
  c.solver.kind = p.solver;
  c.solver.threshold = p.airs_threshold;
  for (int a = 0; a < 3; ++a) c.solver.linv[a] = p.linadv_velocity[a];
  c.gd = make_gas<double>(p.gamma, p.eos, p.isothermal_u);
  c.gf = make_gas<float>(p.gamma, p.eos, p.isothermal_u);
  c.gravity = p.gravity;
  c.grav_soft = p.grav_softening;
  for (int a = 0; a < 3; ++a) {
    c.grav_pos[a] = p.grav_position[a];
    c.grav_acc[a] = p.grav_acceleration[a];
  }
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
  if (p->eos < 0 || p->eos > 1) return set_err(HFV_ERR_INVALID, "unknown equation of state");
  if (!(p->gamma > 1.)) return set_err(HFV_ERR_INVALID, "gamma must be > 1");
  if (!(p->t_end > 0.)) return set_err(HFV_ERR_INVALID, "t_end must be > 0");
  if (p->gravity < HFV_GRAV_NONE || p->gravity > HFV_GRAV_KEPLERIAN) return set_err(HFV_ERR_INVALID, "unknown gravity");
  return HFV_OK;
}

SimEngine fresh_engine(const hfv_params& p) {
  // Engine::new (engine.rs:104-149)
  SimEngine E;
  std::memset(&E, 0, sizeof(E));
  E.time_base = p.t_end / (double)HFV_MAX_NR_TIMESTEPS;
  E.time_base_inv = 1. / E.time_base;
  E.t_end = p.t_end;
  E.t_current = 0.;
  uint64_t between = p.dt_snap > 0. ? sat_u64(p.dt_snap * E.time_base_inv) : HFV_MAX_NR_TIMESTEPS;
  E.between = between ? between : HFV_MAX_NR_TIMESTEPS;
  E.ti_min = ~0ull;
  E.min_bin = 127;
  E.chk = 1;  // the first predict checks the primitives the host writes
  return E;
}

enum { KS_PREDICT = 0, KS_MESH, KS_FLUX, KS_APPLY, KS_GRAD, KS_WAKEUP, KS_LIMIT, KS_DRIFT };

// Device memory of hfv_riemann_batch_noalloc, reserved with the code object on every GPU the library
// runs on (held until the process ends): the Riemann kernel's counters in the first 256 bytes, then the
// staged problems and fluxes of one call, WL[5n] WR[5n] vface[3n] nrm[3n] F[5n], n <= HFV_NOALLOC_MAX_FACES.
// Host code never uses the name `scratch` as an address (on the host it is only a stand-in for the
// symbol): each device's address comes from hipGetSymbolAddress, taken once per device.
constexpr int kScratchHead = 32;  // doubles: the counters' 256 bytes
static_assert(sizeof(DevCounters) <= kScratchHead * sizeof(double), "counters fit the scratch head");
static __device__ __align__(256) double scratch[kScratchHead + 21 * (size_t)HFV_NOALLOC_MAX_FACES];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// The scratch array's address on the current device (looked up on the first call on that device).
int scratch_of_current_device(double** out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) return set_err(HFV_ERR_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char b[256];
      snprintf(b, sizeof(b), "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
      return set_err(HFV_ERR_HIP, b);
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return HFV_OK;
}

}  // namespace

struct hfv_ctx {
  int device = 0;
  hipStream_t stream{};
  int nslots = 0;
  bool batch = false;  // created by hfv_batch_create (sim ids per cell)
  std::vector<hfv_params> params;
  std::vector<int> loaded;
  std::vector<SimDesc> desc;           // host copy
  std::vector<int64_t> ccap, fcap, kcap, cbase, fbase, kbase;
  std::vector<int64_t> nused, fused, kused;
  int64_t n = 0, nf = 0, nconn = 0;    // capacities (totals)
  Cells C{};
  Faces F{};
  // block tables of the dense cell / face kernels (see hfv_kernels.hpp): chunk ranges per slot, the
  // table (position b of the grid holds chunk perm(b)), its host copy and the per-chunk partials
  std::vector<int64_t> cchunk0, fchunk0;
  int64_t ncchunk = 0, nfchunk = 0, ncpos = 0, nfpos = 0;
  // launched grid sizes: positions holding the chunks of live slots (loaded, and in a batch not finished),
  // rounded up to 8; the block tables are rebuilt without finished slots as they finish
  int64_t ncg = 8, nfg = 8, live_chunks = 0;
  std::vector<BlockJob> hcjobs, hfjobs;
  BlockJob *dcjobs = nullptr, *dfjobs = nullptr;
  Partials P{};
  int2* fall = nullptr;      // exact-solver fallback queue (AIRS)
  double* gbuf = nullptr;    // gravity solver output [3n], allocated on first use
  V2Bufs v2{};               // 2D device rebuild (allocated when a 2D moving simulation is loaded)
  bool v2_alloc = false;
  bool use_static = true;    // precomputed geometry for fixed meshes (hfv_set_geometry_mode)
  std::vector<int> v_zero;   // per slot: every particle velocity is +0 (fixed meshes keep it so)
  std::vector<int> has_stage;  // per slot: some face chunk has a staging list
  std::vector<int> has_gstage;  // per slot: some cell chunk has a neighbour staging list
  // per slot, set by hfv_autotune / hfv_set_tuning: LDS-staged flux tiling allowed (1) or not (0);
  // precomputed gradient geometry and LDS-staged gradient neighbours on (1), off (0) or by dimension
  // (-1: on for 3D)
  std::vector<int> tune_stage, tune_sgrad, tune_gstage;
  bool any_static = false, any_dynamic = false, any_static_grad = false, any_dynamic_grad = false;
  SimDesc* ddesc = nullptr;
  SimEngine* deng = nullptr;
  SimEngine* heng = nullptr;  // pinned mirror
  DevCounters* dcnt = nullptr;
  DevCounters* hcnt = nullptr;  // pinned
  std::vector<void*> bufs;
  int precision = HFV_FP64;
  int solver = -1;  // uniform solver of the loaded simulations, or -1
  bool any_moving1d = false;
  bool any_vmesh = false;   // a 2D or 3D moving mesh rebuilt on the device
  bool any_vmesh2 = false;  // ... a 2D one
  bool any_vmesh3 = false;  // ... a 3D one
  bool any_moving = false;
  hipGraph_t graph{};
  hipGraphExec_t gexec{};
  bool graph_ok = false;
  bool use_graph = true;
  int flux_fpt = 1;  // face chunks per flux block (hfv_set_launch_shape, hfv_autotune)
  double tune_fpt_ns[3] = {-1., -1., -1.};  // the tuner's cost per cell update with 1, 2, 4
  bool fr_dirty = true;  // rates and flux records must be rebuilt (k_pack_fr) before the next predict/flux
  hfv_timestep_info last_info{};
  // kernel timing (hfv_profile_enable)
  bool prof = false;
  // a ring of event pairs per kernel slot: every timed step records into its own ring entry, and an
  // entry is read back (waiting for it if needed) before it is recorded again
  static constexpr int PROF_RING = 64;
  hipEvent_t ev0[PROF_RING][HFV_NUM_KERNEL_SLOTS]{}, ev1[PROF_RING][HFV_NUM_KERNEL_SLOTS]{};
  bool pending[PROF_RING][HFV_NUM_KERNEL_SLOTS]{};
  int ring = 0;
  hfv_kernel_times times{};

  Batch batch_args() const {
    Batch B;
    B.C = C;
    B.F = F;
    B.sims = ddesc;
    B.eng = deng;
    B.nsims = nslots;
    B.cjobs = dcjobs;
    B.fjobs = dfjobs;
    B.P = P;
    B.cnt = dcnt;
    B.fall = fall;
    return B;
  }
};

namespace {

// Cells a slot occupies in HBM.  A batch slot carries one permanent empty cell after its capacity: its
// CSR range runs from the end of the slot's used entries to the next slot's first entry, so a slot
// never writes another slot's offsets and no real cell's range covers unused CSR entries.
inline int64_t icap(const hfv_ctx* c, int s) { return c->ccap[s] + (c->batch ? 1 : 0); }

template <class T>
T* dalloc(hfv_ctx* c, int64_t count) {
  void* p = nullptr;
  if (count <= 0) count = 1;
  if (hipMalloc(&p, sizeof(T) * (size_t)count) != hipSuccess) return nullptr;
  c->bufs.push_back(p);
  return (T*)p;
}

void free_all(hfv_ctx* c) {
  for (void* p : c->bufs) (void)hipFree(p);
  c->bufs.clear();
  c->v2 = V2Bufs{};
  c->v2_alloc = false;
  c->gbuf = nullptr;
}

void drop_graph(hfv_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_ok = false;
}

// Buffers of the 2D device rebuild for the whole capacity (hfv_voronoi2d.hpp), on first need.
int ensure_v2(hfv_ctx* c) {
  if (c->v2_alloc) return HFV_OK;
  const int64_t ns = c->n + c->nslots + 1;
  V2Bufs v;
  v.gcnt = dalloc<int32_t>(c, ns);
  v.gfill = dalloc<int32_t>(c, ns);
  v.femit = dalloc<int32_t>(c, ns);
  v.ccnt = dalloc<int32_t>(c, ns);
  v.gord = dalloc<int32_t>(c, c->n);
  v.gkey = dalloc<int32_t>(c, c->n);
  v.vst = dalloc<int32_t>(c, c->n);
  v.cfill = dalloc<int32_t>(c, c->n);
  v.vf = dalloc<V2Face>(c, c->nf);
  v.part = dalloc<int32_t>(c, c->ncchunk);
  if (!v.gcnt || !v.gfill || !v.femit || !v.ccnt || !v.gord || !v.gkey || !v.vst || !v.cfill || !v.vf || !v.part)
    return set_err(HFV_ERR_HIP, "out of device memory (2D rebuild buffers)");
  c->v2 = v;
  c->v2_alloc = true;
  drop_graph(c);  // a captured step holds the old buffer set
  return HFV_OK;
}

// The host tessellator's bin grid for n generators in a 2D or 3D box (hydro_rs_amd/csrc/host/voronoi.cpp,
// Grid::build), evaluated with the same operations.
int v_grid(const hfv_params& p, int64_t n, SimDesc& D) {
  double vol = 1;
  for (int a = 0; a < p.dim; ++a) vol *= p.box[a];
  const double per_gen = vol / (double)std::max<int64_t>(n, 1);
  const double spacing = p.dim == 2 ? std::sqrt(per_gen) : std::cbrt(per_gen);
  for (int a = 0; a < 3; ++a) {
    if (a < p.dim) {
      D.vnb[a] = std::max(1, (int)std::floor(p.box[a] / spacing));
      D.vw[a] = p.box[a] / D.vnb[a];
    } else {
      D.vnb[a] = 1;
      D.vw[a] = 1;
    }
  }
  if ((int64_t)D.vnb[0] * D.vnb[1] * D.vnb[2] > n) return set_err(HFV_ERR_INVALID, "internal: more bins than generators");
  return HFV_OK;
}

// Block-cyclic placement of live chunks on the 8 XCDs (the XCD that runs grid position b is b % 8):
// groups of G consecutive chunks go to the XCDs in turn, so that each XCD's L2 sees runs of neighbouring
// chunks (their gathers overlap) while every XCD gets a share of every part of the work.  G about L/64
// as a power of two, at most 8 in a batch (many simulations of unequal work: balance) and 32 for a
// single simulation (one large mesh: locality).  Measured: ensemble 1.47 -> 1.37 ms per launch against
// whole-simulation ranges per XCD; KH 128^3 2.86 -> 2.81 ms per step.
#ifndef HFV_GROUP_BATCH
#define HFV_GROUP_BATCH 8
#endif
#ifndef HFV_GROUP_SINGLE
#define HFV_GROUP_SINGLE 32
#endif
constexpr int64_t kMaxGroup = HFV_GROUP_BATCH > HFV_GROUP_SINGLE ? HFV_GROUP_BATCH : HFV_GROUP_SINGLE;
int64_t group_size(const hfv_ctx* c, int64_t live) {
  const int64_t cap = c->batch ? HFV_GROUP_BATCH : HFV_GROUP_SINGLE;
  int64_t g = 1;
  while (2 * g <= cap && 64 * (2 * g) <= live) g *= 2;
  return g;
}
// Launched positions for `live` chunks in groups of g: whole rounds of 8 groups.
int64_t grid_size(int64_t live, int64_t g) {
  return std::max<int64_t>(8 * g, (live + 8 * g - 1) / (8 * g) * 8 * g);
}
// Position of live chunk j: group j / g goes to XCD (j / g) % 8, in round j / (8 g).
int64_t grid_position(int64_t j, int64_t g) { return 8 * ((j / g / 8) * g + j % g) + (j / g) % 8; }

// Faces the face kernels cover: the mesh's, or the whole capacity where the device rebuild changes the
// face count (the faces past the current mesh have no cells and are skipped).
int64_t face_range(const hfv_ctx* c, int s) { return c->desc[s].vmesh ? c->fcap[s] : c->fused[s]; }

// Rebuild the block tables from the live slots (loaded, and in a batch not finished: a finished slot's
// blocks would only exit) and upload them with the simulation descriptors.  Chunk ids (partials) stay
// global.
int write_jobs(hfv_ctx* c) {
  const int S = c->nslots;
  auto live = [&](int s) { return c->loaded[s] && !(c->batch && c->heng[s].done); };
  int64_t lc = 0, lf = 0;
  for (int s = 0; s < S; ++s) {
    if (!live(s)) continue;
    lc += (c->nused[s] + BLOCK - 1) / BLOCK;
    lf += (face_range(c, s) + BLOCK - 1) / BLOCK;
  }
  const int64_t gc = group_size(c, lc), gf = group_size(c, lf);
  const int64_t ncg = grid_size(lc, gc), nfg = grid_size(lf, gf);
  if (ncg != c->ncg || nfg != c->nfg) drop_graph(c);  // the step graph holds the grid sizes
  c->ncg = ncg;
  c->nfg = nfg;
  c->live_chunks = lc;
  c->hcjobs.assign(c->ncpos, BlockJob{-1, 0, 0, 0});
  c->hfjobs.assign(c->nfpos, BlockJob{-1, 0, 0, 0});
  int64_t jc = 0, jf = 0;  // running index over the live chunks
  for (int s = 0; s < S; ++s) {
    SimDesc& D = c->desc[s];
    const int64_t nc = c->loaded[s] ? (c->nused[s] + BLOCK - 1) / BLOCK : 0;
    const int64_t nfc = c->loaded[s] ? (face_range(c, s) + BLOCK - 1) / BLOCK : 0;
    D.cchunk0 = (int32_t)c->cchunk0[s];
    D.cchunk1 = (int32_t)(c->cchunk0[s] + nc);
    D.fchunk0 = (int32_t)c->fchunk0[s];
    D.fchunk1 = (int32_t)(c->fchunk0[s] + nfc);
    if (!live(s)) continue;
    for (int64_t j = 0; j < nc; ++j, ++jc) {
      const int64_t k = c->cchunk0[s] + j;
      const int64_t lo = c->cbase[s] + j * BLOCK;
      const int64_t hi = std::min<int64_t>(lo + BLOCK, c->cbase[s] + c->nused[s]);
      c->hcjobs[grid_position(jc, gc)] = BlockJob{s, (int32_t)lo, (int32_t)hi, (int32_t)k};
    }
    for (int64_t j = 0; j < nfc; ++j, ++jf) {
      const int64_t k = c->fchunk0[s] + j;
      const int64_t lo = c->fbase[s] + j * BLOCK;
      const int64_t hi = std::min<int64_t>(lo + BLOCK, c->fbase[s] + face_range(c, s));
      c->hfjobs[grid_position(jf, gf)] = BlockJob{s, (int32_t)lo, (int32_t)hi, (int32_t)k};
    }
  }
  hipStream_t st = c->stream;
  HIP_OK(hipMemcpyAsync(c->dcjobs, c->hcjobs.data(), sizeof(BlockJob) * c->ncpos, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->dfjobs, c->hfjobs.data(), sizeof(BlockJob) * c->nfpos, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->ddesc, c->desc.data(), sizeof(SimDesc) * S, hipMemcpyHostToDevice, st));
  HIP_OK(hipStreamSynchronize(st));
  return HFV_OK;
}

// Allocate every buffer for the slot capacities (cells zeroed, faces' cell indices -1).
int alloc_batch(hfv_ctx* c) {
  free_all(c);
  drop_graph(c);
  int64_t n = 0, nf = 0, nk = 0;
  int S = c->nslots;
  c->cbase.assign(S, 0);
  c->fbase.assign(S, 0);
  c->kbase.assign(S, 0);
  for (int s = 0; s < S; ++s) {
    c->cbase[s] = n;
    c->fbase[s] = nf;
    c->kbase[s] = nk;
    n += icap(c, s);
    nf += c->fcap[s];
    nk += c->kcap[s];
  }
  if (n >= (1ll << 29) || nf >= (1ll << 29) || nk >= (1ll << 30)) return set_err(HFV_ERR_INVALID, "batch too large for 32-bit indices");
  c->n = n;
  c->nf = nf;
  c->nconn = nk;
  Cells& C = c->C;
  C.n = n;
  C.prim = dalloc<double>(c, 5 * n);
  C.grad = dalloc<double>(c, 15 * n);
  C.gc = dalloc<double>(c, 3 * n);
  C.extrap = dalloc<double>(c, 5 * n);
  C.cons = dalloc<double>(c, 5 * n);
  C.facc = dalloc<double>(c, 5 * n);
  C.vol = dalloc<double>(c, n);
  C.loc = dalloc<double>(c, 3 * n);
  C.cen = dalloc<double>(c, 3 * n);
  C.v = dalloc<double>(c, 3 * n);
  C.vrel = dalloc<double>(c, 3 * n);
  C.maor = dalloc<double>(c, n);
  C.msv = dalloc<double>(c, n);
  C.dt = dalloc<double>(c, n);
  C.srad = dalloc<double>(c, n);
  C.pc = dalloc<double>(c, 8 * n);
  C.tex = dalloc<double>(c, 5 * n);
  C.tiex = dalloc<uint64_t>(c, n);
  C.bin = dalloc<int8_t>(c, n);
  C.wake = dalloc<int8_t>(c, n);
  C.act = dalloc<uint8_t>(c, n);
  Faces& F = c->F;
  F.nf = nf;
  F.area = dalloc<double>(c, nf);
  F.nrm = dalloc<double>(c, 3 * nf);
  F.fcen = dalloc<double>(c, 3 * nf);
  F.rec = dalloc<double>(c, 8 * nf);
  F.left = dalloc<int32_t>(c, nf);
  F.right = dalloc<int32_t>(c, nf);
  F.shc = dalloc<uint8_t>(c, nf);
  F.off = dalloc<int32_t>(c, n + 1);
  F.conn = dalloc<int32_t>(c, nk);
  F.nbr = dalloc<int32_t>(c, nk);
  F.cur = dalloc<uint8_t>(c, nf);
  // chunks: a slot's cell capacity and face capacity in blocks of BLOCK items
  c->cchunk0.assign(S, 0);
  c->fchunk0.assign(S, 0);
  int64_t ncc = 0, nfc = 0;
  for (int s = 0; s < S; ++s) {
    c->cchunk0[s] = ncc;
    c->fchunk0[s] = nfc;
    ncc += chunks_of(c->ccap[s]);
    nfc += chunks_of(c->fcap[s]);
  }
  c->ncchunk = ncc;
  c->nfchunk = nfc;
  // room for the largest launched grid (all chunks live, groups of kMaxGroup)
  c->ncpos = grid_size(ncc, kMaxGroup);
  c->nfpos = grid_size(nfc, kMaxGroup);
  c->dcjobs = dalloc<BlockJob>(c, c->ncpos);
  c->dfjobs = dalloc<BlockJob>(c, c->nfpos);
  c->P.act = dalloc<uint32_t>(c, ncc);
  c->P.face = dalloc<uint32_t>(c, nfc);
  c->P.ti = dalloc<uint64_t>(c, ncc);
  c->P.cflag = dalloc<uint32_t>(c, ncc);
  F.nch = dalloc<int32_t>(c, ncc * HFV_NCH_MAX);
  F.nchn = dalloc<uint8_t>(c, ncc);
  c->fall = dalloc<int2>(c, nf);
  C.fr = dalloc<double>(c, (int64_t)HFV_FR_W * n);
  C.dtq = dalloc<double>(c, n);
  F.scell = dalloc<int32_t>(c, nfc * HFV_STAGE_MAX);
  F.scnt = dalloc<uint8_t>(c, nfc);
  F.floc = dalloc<uint8_t>(c, 2 * nf);
  F.gcell = dalloc<int32_t>(c, ncc * HFV_GSTAGE_MAX);
  F.gcnt = dalloc<uint16_t>(c, ncc);
  F.nloc = dalloc<uint16_t>(c, nk);
  C.minv = dalloc<double>(c, 9 * n);
  F.eg = dalloc<double>(c, 8 * nk);
  F.fs = dalloc<double>(c, 4 * nf);
  c->v_zero.assign(S, 1);
  c->has_stage.assign(S, 0);
  c->has_gstage.assign(S, 0);
  c->tune_stage.assign(S, -1);
  c->tune_sgrad.assign(S, -1);
  c->tune_gstage.assign(S, -1);
  c->ddesc = dalloc<SimDesc>(c, S);
  c->deng = dalloc<SimEngine>(c, S);
  for (void* p : c->bufs)
    if (!p) return set_err(HFV_ERR_HIP, "hipMalloc failed");
  hipStream_t st = c->stream;
  const size_t sd = sizeof(double) * (size_t)n;
  double* dbl[] = {C.prim, C.grad, C.gc, C.extrap, C.cons, C.facc, C.vol, C.loc, C.cen, C.v, C.vrel, C.maor, C.msv, C.dt, C.srad, C.tex};
  int mult[] = {5, 15, 3, HFV_EX_W, 5, 5, 1, 3, 3, 3, 3, 1, 1, 1, 1, 5};
  for (int k = 0; k < 16; ++k) HIP_OK(hipMemsetAsync(dbl[k], 0, sd * mult[k], st));
  HIP_OK(hipMemsetAsync(C.pc, 0, sd * 8, st));
  HIP_OK(hipMemsetAsync(C.tiex, 0, sizeof(uint64_t) * (size_t)n, st));
  HIP_OK(hipMemsetAsync(C.fr, 0, sd * HFV_FR_W, st));
  HIP_OK(hipMemsetAsync(F.scnt, 0, (size_t)std::max<int64_t>(nfc, 1), st));
  HIP_OK(hipMemsetAsync(F.floc, 0, 2 * (size_t)std::max<int64_t>(nf, 1), st));
  HIP_OK(hipMemsetAsync(F.gcnt, 0, sizeof(uint16_t) * (size_t)std::max<int64_t>(ncc, 1), st));
  HIP_OK(hipMemsetAsync(C.bin, 0, (size_t)n, st));
  HIP_OK(hipMemsetAsync(C.wake, 0, (size_t)n, st));
  HIP_OK(hipMemsetAsync(C.act, 0, (size_t)n, st));
  HIP_OK(hipMemsetAsync(F.cur, 0, (size_t)nf, st));
  HIP_OK(hipMemsetAsync(c->P.act, 0, sizeof(uint32_t) * (size_t)ncc, st));
  HIP_OK(hipMemsetAsync(c->P.face, 0, sizeof(uint32_t) * (size_t)nfc, st));
  HIP_OK(hipMemsetAsync(c->P.ti, 0xFF, sizeof(uint64_t) * (size_t)ncc, st));
  HIP_OK(hipMemsetAsync(c->P.cflag, 0, sizeof(uint32_t) * (size_t)std::max<int64_t>(ncc, 1), st));
  HIP_OK(hipMemsetAsync(F.nchn, 0xFF, (size_t)std::max<int64_t>(ncc, 1), st));
  HIP_OK(hipMemsetAsync(F.rec, 0, sizeof(double) * 8 * (size_t)nf, st));
  HIP_OK(hipMemsetAsync(F.left, 0xFF, sizeof(int32_t) * (size_t)nf, st));
  HIP_OK(hipMemsetAsync(F.right, 0xFF, sizeof(int32_t) * (size_t)nf, st));
  HIP_OK(hipMemsetAsync(F.off, 0, sizeof(int32_t) * (size_t)(n + 1), st));
  HIP_OK(hipMemsetAsync(F.area, 0, sizeof(double) * (size_t)nf, st));
  HIP_OK(hipMemsetAsync(F.nrm, 0, sizeof(double) * 3 * (size_t)nf, st));
  HIP_OK(hipMemsetAsync(F.fcen, 0, sizeof(double) * 3 * (size_t)nf, st));
  HIP_OK(hipMemsetAsync(F.shc, 0, (size_t)nf, st));
  // engines: all empty (done)
  for (int s = 0; s < S; ++s) {
    std::memset(&c->heng[s], 0, sizeof(SimEngine));
    c->heng[s].done = 1;
  }
  HIP_OK(hipMemcpyAsync(c->deng, c->heng, sizeof(SimEngine) * S, hipMemcpyHostToDevice, st));
  c->desc.assign(S, SimDesc{});
  for (int s = 0; s < S; ++s) {
    std::memset(&c->desc[s], 0, sizeof(SimDesc));
    c->desc[s].cell_base = c->cbase[s];
    c->desc[s].face_base = c->fbase[s];
  }
  HIP_OK(hipMemcpyAsync(c->ddesc, c->desc.data(), sizeof(SimDesc) * S, hipMemcpyHostToDevice, st));
  // CSR offsets: every cell of an empty slot starts at the slot's first CSR entry
  std::vector<int32_t> off(n + 1, 0);
  for (int s = 0; s < S; ++s)
    for (int64_t i = 0; i < icap(c, s); ++i) off[c->cbase[s] + i] = (int32_t)c->kbase[s];
  off[n] = (int32_t)nk;
  HIP_OK(hipMemcpyAsync(F.off, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemsetAsync(F.conn, 0, sizeof(int32_t) * (size_t)std::max<int64_t>(nk, 1), st));
  HIP_OK(hipMemsetAsync(F.nbr, 0xFF, sizeof(int32_t) * (size_t)std::max<int64_t>(nk, 1), st));
  HIP_OK(hipStreamSynchronize(st));
  c->nused.assign(S, 0);
  c->fused.assign(S, 0);
  c->kused.assign(S, 0);
  c->loaded.assign(S, 0);
  return write_jobs(c);
}

int sync_counters(hfv_ctx* c) {
  HIP_OK(hipMemcpyAsync(c->hcnt, c->dcnt, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(c->heng, c->deng, sizeof(SimEngine) * c->nslots, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  // A batch keeps going when one simulation fails: its error is in that slot's engine info.
  if (c->hcnt->err && !c->batch) {
    char b[128];
    snprintf(b, sizeof(b), "device raised hard error %d", c->hcnt->err);
    return set_err(c->hcnt->err, b);
  }
  return HFV_OK;
}

int push_engines(hfv_ctx* c) {
  HIP_OK(hipMemcpyAsync(c->deng, c->heng, sizeof(SimEngine) * c->nslots, hipMemcpyHostToDevice, c->stream));
  return HFV_OK;
}

void recompute_uniforms(hfv_ctx* c) {
  int sol = -2;
  c->any_moving1d = false;
  c->any_vmesh = false;
  c->any_vmesh2 = false;
  c->any_vmesh3 = false;
  c->any_moving = false;
  int prec = -1;
  for (int s = 0; s < c->nslots; ++s) {
    if (!c->loaded[s]) continue;
    const hfv_params& p = c->params[s];
    if (sol == -2) sol = p.solver;
    else if (sol != p.solver) sol = -1;
    if (p.dim == 1 && p.motion != HFV_FIXED) c->any_moving1d = true;
    if (p.dim >= 2 && p.motion != HFV_FIXED) c->any_vmesh = true;
    if (p.dim == 2 && p.motion != HFV_FIXED) c->any_vmesh2 = true;
    if (p.dim == 3 && p.motion != HFV_FIXED) c->any_vmesh3 = true;
    if (p.motion != HFV_FIXED) c->any_moving = true;
    if (prec < 0) prec = p.precision;
  }
  c->solver = sol < 0 ? -1 : sol;
  c->precision = prec < 0 ? HFV_FP64 : prec;
  c->any_static = c->any_dynamic = c->any_static_grad = c->any_dynamic_grad = false;
  for (int s = 0; s < c->nslots; ++s) {
    if (!c->loaded[s]) continue;
    if (c->desc[s].static_geom) c->any_static = true;
    else c->any_dynamic = true;
    if (c->desc[s].static_grad) c->any_static_grad = true;
    else c->any_dynamic_grad = true;
  }
}

// The flags that decide which kernels a captured step launches.
struct LaunchSet {
  int solver;
  int precision;
  bool moving1d, vmesh, vmesh2, vmesh3, moving, st, dyn, sgrad, dgrad;
  bool operator==(const LaunchSet& o) const {
    return solver == o.solver && precision == o.precision && moving1d == o.moving1d && vmesh == o.vmesh &&
           vmesh2 == o.vmesh2 && vmesh3 == o.vmesh3 &&
           moving == o.moving &&
           st == o.st && dyn == o.dyn && sgrad == o.sgrad && dgrad == o.dgrad;
  }
};
LaunchSet launch_set(const hfv_ctx* c) {
  return LaunchSet{c->solver, c->precision, c->any_moving1d, c->any_vmesh, c->any_vmesh2, c->any_vmesh3, c->any_moving, c->any_static,
                   c->any_dynamic, c->any_static_grad, c->any_dynamic_grad};
}

// Upload a mesh into a slot.  update_mask: see hfv_set_mesh.
int upload_mesh(hfv_ctx* c, int s, const hfv_mesh_arrays* m, const uint8_t* update_mask, bool keep_geometry) {
  int64_t n = m->n_cells, nf = m->n_faces;
  if (n <= 0 || nf < 0) return set_err(HFV_ERR_INVALID, "empty mesh");
  if (n > c->ccap[s] || nf > c->fcap[s]) return set_err(HFV_ERR_INVALID, "mesh exceeds the slot's capacity");
  const hfv_params& p = c->params[s];
  const int64_t cb = c->cbase[s], fb = c->fbase[s], kb = c->kbase[s];
  std::vector<double> area(nf), nrm(3 * nf), fcen(3 * nf);
  std::vector<int32_t> left(nf), right(nf);
  std::vector<uint8_t> shc(nf);
  for (int64_t f = 0; f < nf; ++f) {
    if (m->left[f] < 0 || m->left[f] >= n || m->right[f] < -1 || m->right[f] >= n)
      return set_err(HFV_ERR_INVALID, "face cell index out of range");
    area[f] = m->area[f];
    left[f] = (int32_t)(cb + m->left[f]);
    right[f] = m->right[f] < 0 ? -1 : (int32_t)(cb + m->right[f]);
    uint8_t code = 0;
    for (int a = 0; a < 3; ++a) {
      nrm[a * nf + f] = m->normal[3 * f + a];
      fcen[a * nf + f] = m->centroid[3 * f + a];
      double sv = m->shift ? m->shift[3 * f + a] : 0.;
      int cc;
      if (sv == 0.) cc = 0;
      else if (sv == p.box[a]) cc = 1;
      else if (sv == -p.box[a]) cc = 2;
      else return set_err(HFV_ERR_INVALID, "face shift is not 0 or +-box");
      code |= (uint8_t)(cc << (2 * a));
    }
    shc[f] = code;
  }
  // CSR in cell order with side bits; offsets global
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) tot += m->count[i];
  if (tot > c->kcap[s]) return set_err(HFV_ERR_INVALID, "csr exceeds the slot's capacity");
  // offsets of this slot's cells only (plus the global end entry for the last slot)
  const int64_t noff = icap(c, s) + (s == c->nslots - 1 ? 1 : 0);
  std::vector<int32_t> off(noff), conn(tot), nbr(tot);
  int64_t acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    off[i] = (int32_t)(kb + acc);
    for (int64_t k = 0; k < m->count[i]; ++k) {
      int64_t idx = m->offset[i] + k;
      if (idx < 0 || idx >= m->n_conn) return set_err(HFV_ERR_INVALID, "csr offset out of range");
      int64_t f = m->conn[idx];
      if (f < 0 || f >= nf) return set_err(HFV_ERR_INVALID, "csr face index out of range");
      int side;
      if (m->left[f] == i) side = 0;
      else if (m->right[f] == i) side = 1;
      else return set_err(HFV_ERR_INVALID, "csr lists a face that does not touch the cell");
      conn[acc + k] = (int32_t)(((fb + f) << 1) | side);
      const int64_t other = side == 0 ? m->right[f] : m->left[f];
      nbr[acc + k] = other < 0 ? -1 : (int32_t)(cb + other);
    }
    acc += m->count[i];
  }
  for (int64_t i = n; i < noff; ++i) off[i] = (int32_t)(kb + acc);  // empty capacity: no faces
  if (p.dim == 1 && p.motion != HFV_FIXED) {
    int per = p.boundary == HFV_PERIODIC;
    if (nf != (per ? n : n + 1))
      return set_err(HFV_ERR_INVALID, "1D moving mesh needs the 1D tessellator's face layout");
  }
  hipStream_t st = c->stream;
  Faces& F = c->F;
  const int64_t NF = c->nf;
  HIP_OK(hipMemcpyAsync(F.area + fb, area.data(), sizeof(double) * nf, hipMemcpyHostToDevice, st));
  for (int a = 0; a < 3; ++a) {
    HIP_OK(hipMemcpyAsync(F.nrm + a * NF + fb, nrm.data() + a * nf, sizeof(double) * nf, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(F.fcen + a * NF + fb, fcen.data() + a * nf, sizeof(double) * nf, hipMemcpyHostToDevice, st));
  }
  HIP_OK(hipMemcpyAsync(F.left + fb, left.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(F.right + fb, right.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(F.shc + fb, shc.data(), nf, hipMemcpyHostToDevice, st));
  // faces beyond this mesh inside the slot: no cells
  if (nf < c->fcap[s]) {
    HIP_OK(hipMemsetAsync(F.left + fb + nf, 0xFF, sizeof(int32_t) * (c->fcap[s] - nf), st));
    HIP_OK(hipMemsetAsync(F.right + fb + nf, 0xFF, sizeof(int32_t) * (c->fcap[s] - nf), st));
  }
  if (cb + noff > c->n + 1 || (s < c->nslots - 1 && cb + noff > c->cbase[s + 1]))
    return set_err(HFV_ERR_INVALID, "internal: slot offsets overlap the next slot");
  HIP_OK(hipMemcpyAsync(F.off + cb, off.data(), sizeof(int32_t) * noff, hipMemcpyHostToDevice, st));
  if (tot) HIP_OK(hipMemcpyAsync(F.conn + kb, conn.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, st));
  if (tot) HIP_OK(hipMemcpyAsync(F.nbr + kb, nbr.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, st));
  // LDS staging lists of the face chunks (the distinct cells of each chunk's faces, ascending)
  const int64_t nch = (nf + BLOCK - 1) / BLOCK;
  std::vector<int32_t> scell(std::max<int64_t>(nch, 1) * HFV_STAGE_MAX, 0);
  std::vector<uint8_t> scnt(std::max<int64_t>(nch, 1), 0), floc(2 * std::max<int64_t>(nf, 1), 0);
  {
    std::vector<int32_t> u;
    u.reserve(2 * BLOCK);
    for (int64_t j = 0; j < nch; ++j) {
      const int64_t f0 = j * BLOCK, f1 = std::min<int64_t>(f0 + BLOCK, nf);
      u.clear();
      for (int64_t f = f0; f < f1; ++f) {
        u.push_back(left[f]);
        if (right[f] >= 0) u.push_back(right[f]);
      }
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      if ((int64_t)u.size() > HFV_STAGE_MAX) continue;
      scnt[j] = (uint8_t)u.size();
      std::copy(u.begin(), u.end(), scell.begin() + j * HFV_STAGE_MAX);
      for (int64_t f = f0; f < f1; ++f) {
        floc[2 * f] = (uint8_t)(std::lower_bound(u.begin(), u.end(), left[f]) - u.begin());
        floc[2 * f + 1] = right[f] < 0 ? 255 : (uint8_t)(std::lower_bound(u.begin(), u.end(), right[f]) - u.begin());
      }
    }
  }
  if (nch) {
    const int64_t k0 = c->fchunk0[s];
    HIP_OK(hipMemcpyAsync(F.scell + k0 * HFV_STAGE_MAX, scell.data(), sizeof(int32_t) * nch * HFV_STAGE_MAX,
                          hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(F.scnt + k0, scnt.data(), (size_t)nch, hipMemcpyHostToDevice, st));
  }
  if (nf) HIP_OK(hipMemcpyAsync(F.floc + 2 * fb, floc.data(), 2 * (size_t)nf, hipMemcpyHostToDevice, st));
  c->has_stage[s] = std::any_of(scnt.begin(), scnt.end(), [](uint8_t k) { return k > 0; }) ? 1 : 0;
  // LDS staging lists of the cell chunks (the distinct neighbour cells of each chunk's CSR entries)
  {
    const int64_t ncch = (n + BLOCK - 1) / BLOCK;
    std::vector<int32_t> gcell(std::max<int64_t>(ncch, 1) * HFV_GSTAGE_MAX, 0);
    std::vector<uint16_t> gcnt(std::max<int64_t>(ncch, 1), 0), nloc(std::max<int64_t>(tot, 1), 0xFFFF);
    std::vector<int32_t> u;
    u.reserve(16 * BLOCK);
    for (int64_t j = 0; j < ncch; ++j) {
      const int64_t i0 = j * BLOCK, i1 = std::min<int64_t>(i0 + BLOCK, n);
      const int64_t e0 = off[i0] - kb, e1 = (i1 < n ? off[i1] : kb + acc) - kb;
      u.clear();
      for (int64_t e = e0; e < e1; ++e)
        if (nbr[e] >= 0) u.push_back(nbr[e]);
      std::sort(u.begin(), u.end());
      u.erase(std::unique(u.begin(), u.end()), u.end());
      if ((int64_t)u.size() > HFV_GSTAGE_MAX) continue;
      gcnt[j] = (uint16_t)u.size();
      std::copy(u.begin(), u.end(), gcell.begin() + j * HFV_GSTAGE_MAX);
      for (int64_t e = e0; e < e1; ++e)
        if (nbr[e] >= 0) nloc[e] = (uint16_t)(std::lower_bound(u.begin(), u.end(), nbr[e]) - u.begin());
    }
    if (ncch) {
      const int64_t k0 = c->cchunk0[s];
      HIP_OK(hipMemcpyAsync(F.gcell + k0 * HFV_GSTAGE_MAX, gcell.data(), sizeof(int32_t) * ncch * HFV_GSTAGE_MAX,
                            hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(F.gcnt + k0, gcnt.data(), sizeof(uint16_t) * ncch, hipMemcpyHostToDevice, st));
    }
    if (tot) HIP_OK(hipMemcpyAsync(F.nloc + kb, nloc.data(), sizeof(uint16_t) * tot, hipMemcpyHostToDevice, st));
    // the other cell chunks each cell chunk's neighbours lie in (k_apply's idle test); meshes that change
    // on the device (2D moving) keep 0xFF: their chunks always run
    {
      const int64_t cap_ch = chunks_of(c->ccap[s]);
      std::vector<int32_t> nch(cap_ch * HFV_NCH_MAX, 0);
      std::vector<uint8_t> nchn(cap_ch, 0xFF);
      const bool vmesh = p.dim >= 2 && p.motion != HFV_FIXED;
      for (int64_t j = 0; j < ncch && !vmesh; ++j) {
        const int64_t i0 = j * BLOCK, i1 = std::min<int64_t>(i0 + BLOCK, n);
        const int64_t e0 = off[i0] - kb, e1 = (i1 < n ? off[i1] : kb + acc) - kb;
        const int32_t own = (int32_t)(c->cchunk0[s] + j);
        u.clear();
        for (int64_t e = e0; e < e1; ++e)
          if (nbr[e] >= 0) {
            const int32_t ch = (int32_t)(c->cchunk0[s] + (nbr[e] - cb) / BLOCK);
            if (ch != own) u.push_back(ch);
          }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if ((int64_t)u.size() > HFV_NCH_MAX) continue;
        nchn[j] = (uint8_t)u.size();
        std::copy(u.begin(), u.end(), nch.begin() + j * HFV_NCH_MAX);
      }
      const int64_t k0 = c->cchunk0[s];
      HIP_OK(hipMemcpyAsync(F.nch + k0 * HFV_NCH_MAX, nch.data(), sizeof(int32_t) * cap_ch * HFV_NCH_MAX,
                            hipMemcpyHostToDevice, st));
      HIP_OK(hipMemcpyAsync(F.nchn + k0, nchn.data(), (size_t)cap_ch, hipMemcpyHostToDevice, st));
    }
    c->has_gstage[s] = std::any_of(gcnt.begin(), gcnt.end(), [](uint16_t k) { return k > 0; }) ? 1 : 0;
    HIP_OK(hipStreamSynchronize(st));
  }
  // Particle::update_geometry (part.rs:151-158) for the selected cells
  const int64_t N = c->n;
  std::vector<double> vol(n), cen(3 * n), srad(n);
  if (keep_geometry && update_mask) {
    std::vector<double> ov(n), oc(3 * n), os(n);
    HIP_OK(hipMemcpyAsync(ov.data(), c->C.vol + cb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    for (int a = 0; a < 3; ++a)
      HIP_OK(hipMemcpyAsync(oc.data() + a * n, c->C.cen + a * N + cb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(os.data(), c->C.srad + cb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) {
      bool u = update_mask[i];
      vol[i] = u ? m->volume[i] : ov[i];
      srad[i] = u ? m->search_radius[i] : os[i];
      for (int a = 0; a < 3; ++a) cen[a * n + i] = u ? m->cell_centroid[3 * i + a] : oc[a * n + i];
    }
  } else {
    for (int64_t i = 0; i < n; ++i) {
      vol[i] = m->volume[i];
      srad[i] = m->search_radius[i];
      for (int a = 0; a < 3; ++a) cen[a * n + i] = m->cell_centroid[3 * i + a];
    }
  }
  HIP_OK(hipMemcpyAsync(c->C.vol + cb, vol.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  for (int a = 0; a < 3; ++a)
    HIP_OK(hipMemcpyAsync(c->C.cen + a * N + cb, cen.data() + a * n, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->C.srad + cb, srad.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIP_OK(hipStreamSynchronize(st));
  c->nused[s] = n;
  c->fused[s] = nf;
  c->kused[s] = tot;
  SimDesc& D = c->desc[s];
  D.cfg = make_cfg(p);
  D.cell_base = cb;
  D.n = n;
  D.face_base = fb;
  D.nf = nf;
  D.moving1d = p.dim == 1 && p.motion != HFV_FIXED;
  D.periodic = p.boundary == HFV_PERIODIC;
  D.vmesh = p.dim >= 2 && p.motion != HFV_FIXED ? p.dim : 0;
  D.conn_base = c->kbase[s];
  D.conn_cap = c->kcap[s];
  D.face_cap = c->fcap[s];
  D.noff = (int)(icap(c, s) + (s == c->nslots - 1 ? 1 : 0));
  if (D.vmesh) {
    TRY(v_grid(p, n, D));
    TRY(ensure_v2(c));
    c->heng[s].nf_mesh = (int)nf;
    HIP_OK(hipMemcpyAsync(&c->deng[s].nf_mesh, &c->heng[s].nf_mesh, sizeof(int), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
  }
  if (c->loaded[s]) TRY(write_jobs(c));  // a live slot's mesh changed (moving mesh)
  // invariant: every slot's first CSR offset is its CSR base (this slot and the next one)
  for (int t = s; t <= std::min(s + 1, c->nslots - 1); ++t) {
    int32_t first = -1;
    HIP_OK(hipMemcpyAsync(&first, F.off + c->cbase[t], sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (first != (int32_t)c->kbase[t]) return set_err(HFV_ERR_INVALID, "internal: CSR offsets of a slot were overwritten");
  }
  HIP_OK(hipStreamSynchronize(st));
  return HFV_OK;
}

// Reset a slot's per-cell state (Particle::default) and its engine; set the cell -> slot map.
int reset_slot(hfv_ctx* c, int s) {
  hipStream_t st = c->stream;
  const int64_t N = c->n, cb = c->cbase[s], cap = icap(c, s), n = c->nused[s];
  Cells& C = c->C;
  struct {
    double* p;
    int w;
  } fl[] = {{C.prim, 5}, {C.gc, 3}, {C.cons, 5}, {C.facc, 5}, {C.loc, 3}, {C.v, 3},
            {C.vrel, 3}, {C.maor, 1}, {C.msv, 1}, {C.dt, 1}};
  for (auto& f : fl)
    for (int k = 0; k < f.w; ++k) HIP_OK(hipMemsetAsync(f.p + k * N + cb, 0, sizeof(double) * cap, st));
  // extrapolations and kept rates: doubles, or floats in FP32-flux contexts (one precision per batch)
  const size_t es = c->params[s].precision == HFV_FP32_FLUX ? sizeof(float) : sizeof(double);
  HIP_OK(hipMemsetAsync((char*)C.extrap + es * HFV_EX_W * cb, 0, es * HFV_EX_W * cap, st));
  for (int k = 0; k < 5; ++k) HIP_OK(hipMemsetAsync((char*)C.tex + es * (k * N + cb), 0, es * cap, st));
  for (int k = 0; k < 15; ++k) HIP_OK(hipMemsetAsync((char*)C.grad + es * (k * N + cb), 0, es * cap, st));
  HIP_OK(hipMemsetAsync(C.tiex + cb, 0, sizeof(uint64_t) * cap, st));  // = the fresh engine's ti_pred
  HIP_OK(hipMemsetAsync(C.bin + cb, 0, cap, st));
  HIP_OK(hipMemsetAsync(C.wake + cb, 0, cap, st));
  HIP_OK(hipMemsetAsync(C.act + cb, 0, cap, st));
  (void)n;
  c->v_zero[s] = 1;
  c->fr_dirty = true;
  c->heng[s] = fresh_engine(c->params[s]);
  c->heng[s].nf_mesh = (int)c->fused[s];
  HIP_OK(hipMemcpyAsync(c->deng + s, &c->heng[s], sizeof(SimEngine), hipMemcpyHostToDevice, st));
  HIP_OK(hipStreamSynchronize(st));
  return HFV_OK;
}

// Fixed meshes whose particle velocities are all +0 take the static-geometry kernels: their geometric
// terms are evaluated here once (k_geom_*), after every change of mesh, generators or centroids.
int refresh_geometry(hfv_ctx* c, int s) {
  const hfv_params& p = c->params[s];
  const bool st = c->use_static && p.motion == HFV_FIXED && c->v_zero[s] && c->nused[s] > 0;
  c->desc[s].static_geom = st ? 1 : 0;
  c->desc[s].static_grad = st && (c->tune_sgrad[s] < 0 ? p.dim == 3 : c->tune_sgrad[s] != 0) ? 1 : 0;
  // LDS tiling of the flux kernel: by default in 2D/3D; a 1D chunk's faces and cells are already
  // contiguous, so staging there only adds work (512-member ensemble: 2.56e9 -> 2.63e9 without it)
  c->desc[s].staged = st && c->has_stage[s] && (c->tune_stage[s] < 0 ? p.dim != 1 : c->tune_stage[s] != 0) ? 1 : 0;
  c->desc[s].gstaged =
      st && c->has_gstage[s] && (c->tune_gstage[s] < 0 ? p.dim == 3 : c->tune_gstage[s] != 0) ? 1 : 0;
  c->fr_dirty = true;
  HIP_OK(hipMemcpyAsync(c->ddesc + s, &c->desc[s], sizeof(SimDesc), hipMemcpyHostToDevice, c->stream));
  if (c->nused[s] > 0)  // neighbour primitives/centroids of the gradient sweeps
    hipLaunchKernelGGL(k_pack_pc, dim3(grid_full(c->nused[s])), dim3(BLOCK), 0, c->stream, c->batch_args(),
                       c->cbase[s], c->cbase[s] + c->nused[s]);
  if (st) {
    const int64_t cb = c->cbase[s], n = c->nused[s], fb = c->fbase[s], nf = c->fused[s];
    Batch B = c->batch_args();
    if (c->desc[s].static_grad)
      hipLaunchKernelGGL(k_geom_cells, dim3(grid_full(n)), dim3(BLOCK), 0, c->stream, B, s, cb, cb + n);
    if (nf > 0) hipLaunchKernelGGL(k_geom_faces, dim3(grid_full(nf)), dim3(BLOCK), 0, c->stream, B, s, fb, fb + nf);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipStreamSynchronize(c->stream));
  const LaunchSet prev = launch_set(c);
  recompute_uniforms(c);
  if (!(launch_set(c) == prev)) drop_graph(c);
  return HFV_OK;
}

// FP32-flux contexts defer the engine steps' predicts (SimEngine::ti_pred): before a host write or a
// trait-level phase changes what a cell's rate is made of, its unsummed predict time is summed into its
// extrapolations with the old rate (k_settle; slot s, or every slot for s < 0).
void settle(hfv_ctx* c, int s) {
  if (c->precision != HFV_FP32_FLUX || c->ncg <= 0) return;
  hipLaunchKernelGGL(k_settle, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, c->batch_args(), s);
}

int init_ic(hfv_ctx* c, int s, int64_t n, const double* loc, const double* mass, const double* vel, const double* u_int) {
  if (n != c->nused[s] || !loc || !mass || !vel || !u_int) return set_err(HFV_ERR_INVALID, "bad IC arrays");
  const int64_t N = c->n, cb = c->cbase[s];
  std::vector<double> locs(3 * n), vels(3 * n);
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      locs[a * n + i] = loc[3 * i + a];
      vels[a * n + i] = vel[3 * i + a];
    }
  hipStream_t st = c->stream;
  for (int a = 0; a < 3; ++a)
    HIP_OK(hipMemcpyAsync(c->C.loc + a * N + cb, locs.data() + a * n, sizeof(double) * n, hipMemcpyHostToDevice, st));
  double *dm = nullptr, *dv = nullptr, *du = nullptr;
  HIP_OK(hipMalloc((void**)&dm, sizeof(double) * n));
  HIP_OK(hipMalloc((void**)&dv, sizeof(double) * 3 * n));
  HIP_OK(hipMalloc((void**)&du, sizeof(double) * n));
  HIP_OK(hipMemcpyAsync(dm, mass, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(dv, vels.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(du, u_int, sizeof(double) * n, hipMemcpyHostToDevice, st));
  settle(c, s);
  hipLaunchKernelGGL(k_init_ic, dim3(grid_full(n)), dim3(BLOCK), 0, st, c->batch_args(), s, cb, cb + n, dm, dv, du);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(st));
  (void)hipFree(dm);
  (void)hipFree(dv);
  (void)hipFree(du);
  TRY(sync_counters(c));
  return refresh_geometry(c, s);
}

inline void prof_begin(hfv_ctx* c, int slot) {
  if (c->prof) (void)hipEventRecord(c->ev0[c->ring][slot], c->stream);
}
inline void prof_end(hfv_ctx* c, int slot) {
  if (c->prof) {
    (void)hipEventRecord(c->ev1[c->ring][slot], c->stream);
    c->pending[c->ring][slot] = true;
  }
}
// Adds the elapsed time of ring entry r's recorded pairs; `wait` blocks until they have completed.
void prof_take(hfv_ctx* c, int r, bool wait) {
  for (int k = 0; k < HFV_NUM_KERNEL_SLOTS; ++k)
    if (c->pending[r][k]) {
      if (wait) (void)hipEventSynchronize(c->ev1[r][k]);
      float ms = 0.f;
      hipError_t e = hipEventElapsedTime(&ms, c->ev0[r][k], c->ev1[r][k]);
      if (e == hipErrorNotReady) continue;  // read at a later collect
      if (e == hipSuccess) {
        c->times.ms[k] += ms;
        c->times.launches[k] += 1;
      }
      c->pending[r][k] = false;
    }
}
void prof_collect(hfv_ctx* c) {
  if (!c->prof) return;
  for (int r = 0; r < hfv_ctx::PROF_RING; ++r) prof_take(c, r, true);
}
// Moves to the next ring entry before a timed step, reading it back first if it is still pending.
void prof_next(hfv_ctx* c) {
  c->ring = (c->ring + 1) % hfv_ctx::PROF_RING;
  prof_take(c, c->ring, true);
}

template <class T, bool STATIC, int FPT>
void launch_flux_fpt(hfv_ctx* c, const Batch& B, unsigned g) {
  switch (c->solver) {
    case HFV_HLLC:
      hipLaunchKernelGGL((k_flux<T, SOLVER_HLLC, STATIC, FPT>), dim3(g), dim3(BLOCK), 0, c->stream, B);
      break;
    case HFV_AIRS:
      hipLaunchKernelGGL((k_flux<T, SOLVER_AIRS, STATIC, FPT>), dim3(g), dim3(BLOCK), 0, c->stream, B);
      break;
    case HFV_EXACT:
      hipLaunchKernelGGL((k_flux<T, SOLVER_EXACT, STATIC, FPT>), dim3(g), dim3(BLOCK), 0, c->stream, B);
      break;
    default:
      hipLaunchKernelGGL((k_flux<T, -1, STATIC, FPT>), dim3(g), dim3(BLOCK), 0, c->stream, B);
  }
}
template <class T, bool STATIC>
void launch_flux_ts(hfv_ctx* c, const Batch& B, unsigned g) {
  if (c->flux_fpt == 4) launch_flux_fpt<T, STATIC, 4>(c, B, g);
  else if (c->flux_fpt == 2) launch_flux_fpt<T, STATIC, 2>(c, B, g);
  else launch_flux_fpt<T, STATIC, 1>(c, B, g);
}

// Flux kernels of the step: the static-geometry and the general variant each take the blocks of their
// simulations; AIRS queues exact-solver faces for the fallback pass.
void launch_flux(hfv_ctx* c, const Batch& B) {
  const unsigned g = (unsigned)(c->nfg / c->flux_fpt);  // nfg is a multiple of 8
  const bool f32 = c->precision == HFV_FP32_FLUX;
  if (c->any_static) {
    if (f32) launch_flux_ts<float, true>(c, B, g);
    else launch_flux_ts<double, true>(c, B, g);
  }
  if (c->any_dynamic) {
    if (f32) launch_flux_ts<float, false>(c, B, g);
    else launch_flux_ts<double, false>(c, B, g);
  }
  if (c->solver == HFV_AIRS) {
    if (f32) hipLaunchKernelGGL(k_flux_fallback<float>, dim3(256), dim3(BLOCK), 0, c->stream, B);
    else hipLaunchKernelGGL(k_flux_fallback<double>, dim3(256), dim3(BLOCK), 0, c->stream, B);
  }
}

void launch_grad(hfv_ctx* c, const Batch& B, int do_grad, int do_ts) {
  const dim3 g((unsigned)c->ncg), b(BLOCK);
  if (c->any_static_grad) hipLaunchKernelGGL(k_grad<true>, g, b, 0, c->stream, B, do_grad, do_ts);
  if (c->any_dynamic_grad) hipLaunchKernelGGL(k_grad<false>, g, b, 0, c->stream, B, do_grad, do_ts);
}

// The kernels of one batch step (k_begin .. k_end).  Timing events only outside graph capture.
// Space::volume_calculation on the device for the moving meshes being stepped (1D, 2D and 3D).
void enqueue_rebuild(hfv_ctx* c, const Batch& B) {
  hipStream_t st = c->stream;
  const dim3 gc((unsigned)c->ncg), gf((unsigned)c->nfg), gb(BLOCK);
  if (c->any_moving1d) {
    hipLaunchKernelGGL(k_mesh1d_faces, gf, gb, 0, st, B);
    hipLaunchKernelGGL(k_mesh1d_cells, gc, gb, 0, st, B);
  }
  if (c->any_vmesh) {
    const V2Bufs& V = c->v2;
    const dim3 gs((unsigned)c->nslots), gsb(HFV_SCAN_BLOCK);
    hipLaunchKernelGGL(k_v2_clear, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_bin, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_up<0>, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_mid<0>, gs, gsb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_down<0>, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_place, gc, gb, 0, st, B, V);
    if (c->any_vmesh2) {
      hipLaunchKernelGGL(k_v2_cells<false>, dim3((unsigned)(c->ncg * HFV_V2_SUB)), gb, 0, st, B, V);
      hipLaunchKernelGGL(k_v2_cells<true>, dim3((unsigned)(c->ncg * HFV_V2_SUB)), gb, 0, st, B, V);
    }
    if (c->any_vmesh3) {
      const dim3 g3((unsigned)(c->ncg * BLOCK)), w3(64);
      hipLaunchKernelGGL(k_v3_cells, g3, w3, 0, st, B, V);
    }
    hipLaunchKernelGGL(k_v2_scan_up<1>, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_mid<1>, gs, gsb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_down<1>, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_faces, gf, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_up<2>, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_mid<2>, gs, gsb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_scan_down<2>, gc, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_csr_fill, gf, gb, 0, st, B, V);
    hipLaunchKernelGGL(k_v2_csr_sort, gc, gb, 0, st, B, V);
  }
}

void enqueue_step(hfv_ctx* c, bool timed) {
  Batch B = c->batch_args();
  hipStream_t st = c->stream;
  const unsigned gs = (unsigned)std::max(1, (c->nslots + 63) / 64);
  const dim3 gc((unsigned)c->ncg), gf((unsigned)c->nfg), gb(BLOCK);
  const bool t = timed && c->prof;
  if (t) prof_next(c);
  hipLaunchKernelGGL(k_begin, dim3(gs), dim3(64), 0, st, B);
  if (c->any_moving) {
    if (t) prof_begin(c, KS_DRIFT);
    hipLaunchKernelGGL(k_drift, gc, gb, 0, st, B);
    if (t) prof_end(c, KS_DRIFT);
  }
  if (t) prof_begin(c, KS_PREDICT);
  hipLaunchKernelGGL(k_mask, gc, gb, 0, st, B, 0);
  if (t) prof_end(c, KS_PREDICT);
  if (c->any_moving1d || c->any_vmesh) {
    if (t) prof_begin(c, KS_MESH);
    enqueue_rebuild(c, B);
    if (t) prof_end(c, KS_MESH);
  }
  if (t) prof_begin(c, KS_FLUX);
  launch_flux(c, B);
  if (t) prof_end(c, KS_FLUX);
  if (t) prof_begin(c, KS_APPLY);
  hipLaunchKernelGGL(k_apply, gc, gb, 0, st, B, 7);
  if (t) prof_end(c, KS_APPLY);
  if (t) prof_begin(c, KS_GRAD);
  launch_grad(c, B, 1, 1);
  if (t) prof_end(c, KS_GRAD);
  if (t) prof_begin(c, KS_WAKEUP);
  hipLaunchKernelGGL(k_wakeup, gc, gb, 0, st, B, 1);
  if (t) prof_end(c, KS_WAKEUP);
  if (t) prof_begin(c, KS_LIMIT);
  hipLaunchKernelGGL(k_limit, gc, gb, 0, st, B, 0);
  if (t) prof_end(c, KS_LIMIT);
  hipLaunchKernelGGL(k_end, dim3((unsigned)c->nslots), dim3(HFV_END_BLOCK), 0, st, B);
}

// Rebuild the kept time-extrapolation rates (and the staged simulations' flux records) when the host or a
// phase call changed their inputs (the fused step keeps them current itself).
void pack_fr_if_dirty(hfv_ctx* c) {
  if (c->fr_dirty)
    hipLaunchKernelGGL(k_pack_fr, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, c->batch_args());
  c->fr_dirty = false;
}

int launch_step(hfv_ctx* c) {
  pack_fr_if_dirty(c);
  if (c->prof || !c->use_graph) {
    enqueue_step(c, true);
    HIP_OK(hipGetLastError());
    return HFV_OK;
  }
  if (!c->graph_ok) {
    drop_graph(c);
    HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, false);
    HIP_OK(hipStreamEndCapture(c->stream, &c->graph));
    HIP_OK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    c->graph_ok = true;
  }
  HIP_OK(hipGraphLaunch(c->gexec, c->stream));
  return HFV_OK;
}

// Faces of a slot's current mesh (after sync_counters for device-rebuilt meshes).
int64_t mesh_faces(const hfv_ctx* c, int s) { return c->desc[s].vmesh ? c->heng[s].nf_mesh : c->fused[s]; }

int check_slot(hfv_ctx* c, int s) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  if (s < 0 || s >= c->nslots) return set_err(HFV_ERR_INVALID, "slot out of range");
  HIP_OK(hipSetDevice(c->device));
  return HFV_OK;
}

int require_mesh(hfv_ctx* c) {
  TRY(check_slot(c, 0));
  if (!c->loaded[0]) return set_err(HFV_ERR_INVALID, "mesh not set");
  return HFV_OK;
}

int new_ctx(hfv_ctx** out, int nslots) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(HFV_ERR_NO_DEVICE, "no HIP device");
  hfv_ctx* c = new hfv_ctx();
  c->nslots = nslots;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "hipGetDevice failed");
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "hipStreamCreate failed");
  }
  if (hipMalloc((void**)&c->dcnt, sizeof(DevCounters)) != hipSuccess ||
      hipHostMalloc((void**)&c->hcnt, sizeof(DevCounters), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->heng, sizeof(SimEngine) * (size_t)std::max(1, nslots), hipHostMallocDefault) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "counter allocation failed");
  }
  std::memset(c->hcnt, 0, sizeof(DevCounters));
  if (hipMemcpy(c->dcnt, c->hcnt, sizeof(DevCounters), hipMemcpyHostToDevice) != hipSuccess) {
    delete c;
    return set_err(HFV_ERR_HIP, "counter init failed");
  }
  c->params.assign(nslots, hfv_params{});
  c->loaded.assign(nslots, 0);
  *out = c;
  return HFV_OK;
}

void fill_engine_info(const SimEngine& E, hfv_engine_info* out) {
  out->ts.ti_old = E.ti_old;
  out->ts.ti_current = E.ti_current;
  out->ts.ti_next = E.ti_next;
  out->ts.time_base = E.time_base;
  out->ts.time_base_inv = E.time_base_inv;
  out->t_current = E.t_current;
  out->t_end = E.t_end;
  out->step_count = E.step_count;
  out->cell_updates = E.cell_updates;
  out->face_fluxes = E.face_fluxes;
  out->last_n_active = E.last_active;
  out->last_n_face_active = E.last_face_active;
  out->done = E.done || !(E.t_current < E.t_end);
  out->error_code = E.err;
}

#define STATE_FIELDS(X)                                                                           \
  X(prim, C.prim, 5) X(gc, C.gc, 3) X(cons, C.cons, 5)                                            \
  X(fluxes, C.facc, 5) X(volume, C.vol, 1) X(loc, C.loc, 3) X(centroid, C.cen, 3) X(v, C.v, 3)    \
  X(v_rel, C.vrel, 3) X(maor, C.maor, 1) X(msv, C.msv, 1) X(dt, C.dt, 1)

int copy_state(hfv_ctx* c, int s, hfv_state* st, bool to_device) {
  const int64_t N = c->n, cb = c->cbase[s], n = c->nused[s];
  hipStream_t strm = c->stream;
  hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  if (to_device) settle(c, s);
#define CPF(name, dev, w)                                                                        \
  if (st->name)                                                                                  \
    for (int q = 0; q < (w); ++q) {                                                              \
      double* d = c->dev + q * N + cb;                                                           \
      double* h = st->name + q * n;                                                              \
      HIP_OK(hipMemcpyAsync(to_device ? (void*)d : (void*)h, to_device ? (const void*)h : (const void*)d, \
                            sizeof(double) * n, kind, strm));                                    \
    }
  STATE_FIELDS(CPF)
#undef CPF
  // extrapolations: AoS [N][5] on the device, component-major [5][n] at the boundary
  // (floats in FP32-flux contexts, as the gradients)
  const bool ex32 = c->params[s].precision == HFV_FP32_FLUX;
  std::vector<float> grf;
  if (st->grad) {
    if (ex32) {
      grf.resize(15 * (size_t)n);
      for (int q = 0; q < 15; ++q) {
        float* d = (float*)c->C.grad + (size_t)q * N + cb;
        if (to_device) {
          for (int64_t i = 0; i < n; ++i) grf[(size_t)q * n + i] = (float)st->grad[q * n + i];
          HIP_OK(hipMemcpyAsync(d, grf.data() + (size_t)q * n, sizeof(float) * n, hipMemcpyHostToDevice, strm));
        } else {
          HIP_OK(hipMemcpyAsync(grf.data() + (size_t)q * n, d, sizeof(float) * n, hipMemcpyDeviceToHost, strm));
        }
      }
    } else {
      for (int q = 0; q < 15; ++q) {
        double* d = c->C.grad + (size_t)q * N + cb;
        double* h = st->grad + (size_t)q * n;
        HIP_OK(hipMemcpyAsync(to_device ? (void*)d : (void*)h, to_device ? (const void*)h : (const void*)d,
                              sizeof(double) * n, kind, strm));
      }
    }
  }
  std::vector<double> exa;
  std::vector<float> exf;
  if (st->extrap) {
    exa.resize(HFV_EX_W * (size_t)n);
    if (ex32) exf.resize(HFV_EX_W * (size_t)n);
    void* d = ex32 ? (void*)((float*)c->C.extrap + HFV_EX_W * cb) : (void*)(c->C.extrap + HFV_EX_W * cb);
    const size_t bytes = (ex32 ? sizeof(float) : sizeof(double)) * HFV_EX_W * (size_t)n;
    if (to_device) {
      for (int64_t i = 0; i < n; ++i)
        for (int q = 0; q < 5; ++q) {
          exa[HFV_EX_W * i + q] = st->extrap[q * n + i];
          if (ex32) exf[HFV_EX_W * i + q] = (float)st->extrap[q * n + i];
        }
      HIP_OK(hipMemcpyAsync(d, ex32 ? (const void*)exf.data() : (const void*)exa.data(), bytes, hipMemcpyHostToDevice,
                            strm));
    } else {
      HIP_OK(hipMemcpyAsync(ex32 ? (void*)exf.data() : (void*)exa.data(), d, bytes, hipMemcpyDeviceToHost, strm));
    }
  }
  // FP32-flux reads: the extrapolations due now, summed on the host (k_settle's expression) so that a
  // read leaves the device state as it is
  const bool due = st->extrap && ex32 && !to_device;
  std::vector<uint64_t> tix;
  std::vector<float> txf;
  std::vector<double> rho;
  SimEngine eh;
  if (due) {
    tix.resize(n);
    txf.resize(5 * (size_t)n);
    rho.resize(n);
    HIP_OK(hipMemcpyAsync(tix.data(), c->C.tiex + cb, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, strm));
    for (int q = 0; q < 5; ++q)
      HIP_OK(hipMemcpyAsync(txf.data() + (size_t)q * n, (const float*)c->C.tex + (size_t)q * N + cb, sizeof(float) * n,
                            hipMemcpyDeviceToHost, strm));
    HIP_OK(hipMemcpyAsync(rho.data(), c->C.prim + cb, sizeof(double) * n, hipMemcpyDeviceToHost, strm));
    HIP_OK(hipMemcpyAsync(&eh, c->deng + s, sizeof(SimEngine), hipMemcpyDeviceToHost, strm));
  }
  if (st->timebin)
    HIP_OK(hipMemcpyAsync(to_device ? (void*)(c->C.bin + cb) : (void*)st->timebin,
                          to_device ? (const void*)st->timebin : (const void*)(c->C.bin + cb), (size_t)n, kind, strm));
  HIP_OK(hipStreamSynchronize(strm));
  if (st->grad && ex32 && !to_device)
    for (size_t k = 0; k < grf.size(); ++k) st->grad[k] = (double)grf[k];
  if (st->extrap && !to_device)
    for (int64_t i = 0; i < n; ++i) {
      const double dt = due ? (double)(eh.ti_pred - tix[i]) * eh.time_base : 0.;
      for (int q = 0; q < 5; ++q) {
        double v = ex32 ? (double)exf[HFV_EX_W * i + q] : exa[HFV_EX_W * i + q];
        if (due && rho[i] > 0.) v = (double)(float)(v + (-dt) * (double)txf[(size_t)q * n + i]);
        st->extrap[q * n + i] = v;
      }
    }
  if (to_device) {
    c->fr_dirty = true;
    if (st->v) {  // the static geometry needs every particle velocity to be exactly +0
      int z = 1;
      for (int64_t k = 0; k < 3 * n && z; ++k) {
        uint64_t bits;
        std::memcpy(&bits, st->v + k, sizeof(bits));
        z = bits == 0;
      }
      c->v_zero[s] = z;
    }
    if (st->loc || st->centroid || st->v || st->prim) TRY(refresh_geometry(c, s));
  }
  return HFV_OK;
}

// Trait-level phases operate on slot 0 with the caller's TimestepInfo.
// pred: the step's predict is the engine's own (hfv_step), not a separate hfv_predict call.
int begin_phase_step(hfv_ctx* c, const hfv_timestep_info* info, bool pred) {
  SimEngine& E = c->heng[0];
  HIP_OK(hipMemcpyAsync(&E, c->deng, sizeof(SimEngine), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  E.ti_old = info->ti_old;
  E.ti_current = info->ti_current;
  E.ti_next = info->ti_current;
  E.time_base = info->time_base;
  E.time_base_inv = info->time_base_inv;
  E.force = 1;
  E.done = 0;
  E.no_pred = pred ? 0 : 1;
  HIP_OK(hipMemcpyAsync(c->deng, &E, sizeof(SimEngine), hipMemcpyHostToDevice, c->stream));
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
  TRY(check_params(params));
  if (!out) return set_err(HFV_ERR_INVALID, "out is NULL");
  hfv_ctx* c = nullptr;
  TRY(new_ctx(&c, 1));
  c->params[0] = *params;
  c->batch = false;
  *out = c;
  return HFV_OK;
}

int hfv_destroy(hfv_ctx* c) {
  if (!c) return HFV_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_all(c);
  for (int r = 0; r < hfv_ctx::PROF_RING; ++r)
    for (int k = 0; k < HFV_NUM_KERNEL_SLOTS; ++k) {
      if (c->ev0[r][k]) (void)hipEventDestroy(c->ev0[r][k]);
      if (c->ev1[r][k]) (void)hipEventDestroy(c->ev1[r][k]);
    }
  if (c->dcnt) (void)hipFree(c->dcnt);
  if (c->hcnt) (void)hipHostFree(c->hcnt);
  if (c->heng) (void)hipHostFree(c->heng);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return HFV_OK;
}

int hfv_set_mesh(hfv_ctx* c, const hfv_mesh_arrays* m, const uint8_t* update_mask) {
  TRY(check_slot(c, 0));
  if (!m) return set_err(HFV_ERR_INVALID, "mesh is NULL");
  if (c->batch) return set_err(HFV_ERR_INVALID, "use hfv_batch_load on a batch context");
  bool first = !c->loaded[0];
  if (!first && m->n_cells != c->nused[0]) return set_err(HFV_ERR_INVALID, "cell count differs from the existing state");
  if (first || m->n_faces > c->fcap[0] || m->n_conn > c->kcap[0]) {
    if (!first) {
      // growing the face arrays of a live simulation: keep the cell state across the reallocation
      return set_err(HFV_ERR_INVALID, "mesh has more faces than the context was created for");
    }
    c->ccap.assign(1, m->n_cells);
    // headroom for meshes rebuilt later (moving mesh)
    c->fcap.assign(1, m->n_faces + m->n_faces / 4 + 16);
    int64_t tot = 0;
    for (int64_t i = 0; i < m->n_cells; ++i) tot += m->count[i];
    c->kcap.assign(1, std::max<int64_t>(tot, m->n_conn) + std::max<int64_t>(tot, m->n_conn) / 4 + 32);
    TRY(alloc_batch(c));
  }
  TRY(upload_mesh(c, 0, m, update_mask, !first));
  if (first) {
    c->loaded[0] = 1;
    TRY(write_jobs(c));
    TRY(reset_slot(c, 0));
  } else {
    TRY(refresh_geometry(c, 0));
  }
  recompute_uniforms(c);
  drop_graph(c);
  return HFV_OK;
}

int hfv_init_from_ic(hfv_ctx* c, int64_t n, const double* loc, const double* mass, const double* vel,
                     const double* u_int) {
  TRY(require_mesh(c));
  return init_ic(c, 0, n, loc, mass, vel, u_int);
}

int hfv_set_state(hfv_ctx* c, const hfv_state* st) {
  TRY(require_mesh(c));
  if (!st) return set_err(HFV_ERR_INVALID, "state is NULL");
  return copy_state(c, 0, const_cast<hfv_state*>(st), true);
}

int hfv_get_state(hfv_ctx* c, hfv_state* st) {
  TRY(require_mesh(c));
  if (!st) return set_err(HFV_ERR_INVALID, "state is NULL");
  return copy_state(c, 0, st, false);
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

// ---- FiniteVolumeSolver trait surface (slot 0, caller-provided TimestepInfo)
int hfv_predict(hfv_ctx* c, double dt) {
  TRY(require_mesh(c));
  SimEngine& E = c->heng[0];
  HIP_OK(hipMemcpyAsync(&E, c->deng, sizeof(SimEngine), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  E.dt_pred = dt;
  HIP_OK(hipMemcpyAsync(c->deng, &E, sizeof(SimEngine), hipMemcpyHostToDevice, c->stream));
  pack_fr_if_dirty(c);  // predict reads the kept time-extrapolation rates
  hipLaunchKernelGGL(k_mask, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, c->batch_args(), 1);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  return sync_counters(c);
}

int hfv_compute_fluxes(hfv_ctx* c, const hfv_timestep_info* info) {
  TRY(require_mesh(c));
  if (!info) return set_err(HFV_ERR_INVALID, "info is NULL");
  c->last_info = *info;
  TRY(begin_phase_step(c, info, false));
  Batch B = c->batch_args();
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, c->stream, B);
  hipLaunchKernelGGL(k_mask, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, B, 2);
  c->fr_dirty = true;
  pack_fr_if_dirty(c);
  launch_flux(c, B);
  HIP_OK(hipGetLastError());
  return sync_counters(c);
}

int hfv_get_fluxes(hfv_ctx* c, double* out7f) {
  TRY(require_mesh(c));
  if (!out7f) return set_err(HFV_ERR_INVALID, "out is NULL");
  TRY(sync_counters(c));
  const int64_t NF = c->nf, nf = mesh_faces(c, 0), fb = c->fbase[0], n = c->nused[0], cb = c->cbase[0];
  std::vector<double> aos(8 * nf);
  std::vector<uint8_t> act(n);
  std::vector<int32_t> left(nf), right(nf);
  (void)NF;
  std::vector<float> aosf;
  if (c->precision == HFV_FP32_FLUX) {  // FP32-flux records: 8 floats per face
    aosf.resize(8 * nf);
    HIP_OK(hipMemcpyAsync(aosf.data(), (const float*)c->F.rec + 8 * fb, sizeof(float) * 8 * nf, hipMemcpyDeviceToHost,
                          c->stream));
  } else {
    HIP_OK(hipMemcpyAsync(aos.data(), c->F.rec + 8 * fb, sizeof(double) * 8 * nf, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_OK(hipMemcpyAsync(act.data(), c->C.act + cb, n, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(left.data(), c->F.left + fb, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(right.data(), c->F.right + fb, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int64_t f = 0; f < nf; ++f) {
    // records are current for faces with an active cell; the others are gated off (zero)
    bool valid = act[left[f] - cb] || (right[f] >= 0 && act[right[f] - cb]);
    for (int k = 0; k < 7; ++k)
      out7f[7 * f + k] = !valid ? 0. : aosf.empty() ? aos[8 * f + k] : (double)aosf[8 * f + k];
  }
  return HFV_OK;
}

int hfv_apply_fluxes(hfv_ctx* c) {
  TRY(require_mesh(c));
  settle(c, -1);
  hipLaunchKernelGGL(k_apply, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, c->batch_args(), 1);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  return sync_counters(c);
}

int hfv_conserved_to_primitive(hfv_ctx* c) {
  TRY(require_mesh(c));
  settle(c, -1);
  hipLaunchKernelGGL(k_apply, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, c->batch_args(), 2);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  return sync_counters(c);
}

int hfv_gradients(hfv_ctx* c) {
  TRY(require_mesh(c));
  settle(c, -1);
  hipLaunchKernelGGL(k_apply, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, c->batch_args(), 4 | 8);
  launch_grad(c, c->batch_args(), 1, 0);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  return sync_counters(c);
}

int hfv_timesteps(hfv_ctx* c, uint64_t* ti_next) {
  TRY(require_mesh(c));
  settle(c, -1);
  Batch B = c->batch_args();
  launch_grad(c, B, 0, 1);
  hipLaunchKernelGGL(k_wakeup, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, B, 0);
  hipLaunchKernelGGL(k_limit, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, B, 1);
  hipLaunchKernelGGL(k_end, dim3((unsigned)c->nslots), dim3(HFV_END_BLOCK), 0, c->stream, B);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  TRY(sync_counters(c));
  if (c->heng[0].err) return set_err(c->heng[0].err, "engine error in this step");
  if (ti_next) *ti_next = c->heng[0].ti_next;
  return HFV_OK;
}

int hfv_compute_timesteps(hfv_ctx* c, double* dt_out) {
  TRY(require_mesh(c));
  if (!dt_out) return set_err(HFV_ERR_INVALID, "out is NULL");
  settle(c, -1);
  launch_grad(c, c->batch_args(), 0, 2);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  TRY(sync_counters(c));
  const int64_t n = c->nused[0], cb = c->cbase[0];
  std::vector<uint8_t> act(n);
  HIP_OK(hipMemcpyAsync(dt_out, c->C.dtq + cb, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipMemcpyAsync(act.data(), c->C.act + cb, n, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i)
    if (!act[i]) dt_out[i] = INFINITY;  // euler_eqns_fvs.rs:52
  return HFV_OK;
}

int hfv_apply_timesteps(hfv_ctx* c, const double* dt, uint64_t* ti_next) {
  TRY(require_mesh(c));
  if (!dt) return set_err(HFV_ERR_INVALID, "dt is NULL");
  settle(c, -1);
  const int64_t n = c->nused[0], cb = c->cbase[0];
  HIP_OK(hipMemcpyAsync(c->C.dtq + cb, dt, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  Batch B = c->batch_args();
  hipLaunchKernelGGL(k_ts_apply, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, B);
  hipLaunchKernelGGL(k_wakeup, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, B, 0);
  hipLaunchKernelGGL(k_limit, dim3((unsigned)c->ncg), dim3(BLOCK), 0, c->stream, B, 1);
  hipLaunchKernelGGL(k_end, dim3((unsigned)c->nslots), dim3(HFV_END_BLOCK), 0, c->stream, B);
  HIP_OK(hipGetLastError());
  c->fr_dirty = true;
  TRY(sync_counters(c));
  if (c->heng[0].err) return set_err(c->heng[0].err, "engine error in this step");
  if (ti_next) *ti_next = c->heng[0].ti_next;
  return HFV_OK;
}

// ---- the current mesh (Space's faces and connections, space.rs:76-77, and the cells' geometry)
int hfv_get_mesh(hfv_ctx* c, hfv_mesh_arrays* out) {
  TRY(require_mesh(c));
  if (!out) return set_err(HFV_ERR_INVALID, "out is NULL");
  TRY(sync_counters(c));
  const int64_t n = c->nused[0], nf = mesh_faces(c, 0), cb = c->cbase[0], fb = c->fbase[0], NF = c->nf, N = c->n;
  hipStream_t st = c->stream;
  std::vector<int32_t> off(n + 1);
  HIP_OK(hipMemcpyAsync(off.data(), c->F.off + cb, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  const int64_t nk = off[n] - off[0];
  out->n_cells = n;
  out->n_faces = nf;
  out->n_conn = nk;
  if (!out->area) return HFV_OK;  // sizes only
  if (!out->normal || !out->centroid || !out->shift || !out->left || !out->right || !out->volume ||
      !out->cell_centroid || !out->search_radius || !out->offset || !out->count || !out->conn)
    return set_err(HFV_ERR_INVALID, "an output array is NULL");
  std::vector<double> nrm(3 * nf), fcen(3 * nf), cen(3 * n);
  std::vector<int32_t> left(nf), right(nf), conn(nk);
  std::vector<uint8_t> shc(nf);
  HIP_OK(hipMemcpyAsync((double*)out->area, c->F.area + fb, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
  for (int a = 0; a < 3; ++a) {
    HIP_OK(hipMemcpyAsync(nrm.data() + a * nf, c->F.nrm + a * NF + fb, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(fcen.data() + a * nf, c->F.fcen + a * NF + fb, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(cen.data() + a * n, c->C.cen + a * N + cb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  }
  HIP_OK(hipMemcpyAsync(left.data(), c->F.left + fb, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(right.data(), c->F.right + fb, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(shc.data(), c->F.shc + fb, nf, hipMemcpyDeviceToHost, st));
  if (nk) HIP_OK(hipMemcpyAsync(conn.data(), c->F.conn + off[0], sizeof(int32_t) * nk, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync((double*)out->volume, c->C.vol + cb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync((double*)out->search_radius, c->C.srad + cb, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  const hfv_params& p = c->params[0];
  double* onrm = (double*)out->normal;
  double* ocen = (double*)out->centroid;
  double* osh = (double*)out->shift;
  int64_t* ol = (int64_t*)out->left;
  int64_t* orr = (int64_t*)out->right;
  for (int64_t f = 0; f < nf; ++f) {
    for (int a = 0; a < 3; ++a) {
      onrm[3 * f + a] = nrm[a * nf + f];
      ocen[3 * f + a] = fcen[a * nf + f];
      const int code = (shc[f] >> (2 * a)) & 3;
      osh[3 * f + a] = code == 1 ? p.box[a] : code == 2 ? -p.box[a] : 0.;
    }
    ol[f] = left[f] - cb;
    orr[f] = right[f] < 0 ? -1 : right[f] - cb;
  }
  double* occ = (double*)out->cell_centroid;
  int64_t* ooff = (int64_t*)out->offset;
  int64_t* ocnt = (int64_t*)out->count;
  int64_t* oconn = (int64_t*)out->conn;
  for (int64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) occ[3 * i + a] = cen[a * n + i];
    ooff[i] = off[i] - off[0];
    ocnt[i] = off[i + 1] - off[i];
  }
  for (int64_t k = 0; k < nk; ++k) oconn[k] = (conn[k] >> 1) - fb;
  return HFV_OK;
}

// ---- GravitySolver trait (src/gravity.rs:6-10)
static int gravity_phase(hfv_ctx* c, double* out, bool accelerations) {
  TRY(require_mesh(c));
  if (!out) return set_err(HFV_ERR_INVALID, "out is NULL");
  if (c->params[0].gravity == HFV_GRAV_NONE) return set_err(HFV_ERR_INVALID, "the context has no gravity solver");
  if (!c->gbuf && !(c->gbuf = dalloc<double>(c, 3 * c->n))) return set_err(HFV_ERR_HIP, "out of device memory");
  settle(c, -1);
  const int64_t n = c->nused[0], cb = c->cbase[0];
  const Cfg cfg = make_cfg(c->params[0]);
  const unsigned grid = (unsigned)((n + BLOCK - 1) / BLOCK);
  if (n > 0) {
    if (accelerations)
      hipLaunchKernelGGL(k_grav_acc, dim3(grid), dim3(BLOCK), 0, c->stream, c->batch_args(), cfg, cb, n, c->gbuf);
    else
      hipLaunchKernelGGL(k_grav_dt, dim3(grid), dim3(BLOCK), 0, c->stream, c->batch_args(), cfg, cb, n, c->gbuf);
    HIP_OK(hipGetLastError());
  }
  const int64_t width = accelerations ? 3 : 1;
  HIP_OK(hipMemcpyAsync(out, c->gbuf, sizeof(double) * width * n, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  return HFV_OK;
}

int hfv_gravity_timesteps(hfv_ctx* c, double* dt_out) { return gravity_phase(c, dt_out, false); }

int hfv_gravity_accelerations(hfv_ctx* c, double* a_out) { return gravity_phase(c, a_out, true); }

// ---- fused step with the caller's TimestepInfo
int hfv_step(hfv_ctx* c, const hfv_timestep_info* info, uint64_t* ti_next) {
  TRY(require_mesh(c));
  if (!info) return set_err(HFV_ERR_INVALID, "info is NULL");
  c->last_info = *info;
  TRY(begin_phase_step(c, info, true));
  pack_fr_if_dirty(c);
  enqueue_step(c, true);
  HIP_OK(hipGetLastError());
  TRY(sync_counters(c));
  prof_collect(c);
  if (c->heng[0].err) return set_err(c->heng[0].err, "engine error in this step");
  if (ti_next) *ti_next = c->heng[0].ti_next;
  return HFV_OK;
}

// ---- engine (src/engine.rs:152-268), device-resident
// The device engine loops drift generators (k_drift) and rebuild moving meshes in every dimension.

int hfv_run(hfv_ctx* c, int64_t max_steps, int64_t* steps_done) {
  TRY(require_mesh(c));
  TRY(sync_counters(c));
  const long long s0 = c->heng[0].step_count;
  int64_t done = 0;
  int64_t chunk = 8;
  while (done < max_steps && !c->heng[0].done) {
    int64_t k = std::min<int64_t>(chunk, max_steps - done);
    for (int64_t j = 0; j < k; ++j) TRY(launch_step(c));
    TRY(sync_counters(c));
    prof_collect(c);
    done = c->heng[0].step_count - s0;
    if (c->heng[0].err) return set_err(c->heng[0].err, "engine error");
    chunk = std::min<int64_t>(chunk * 2, 256);
  }
  if (steps_done) *steps_done = done;
  return HFV_OK;
}

int hfv_get_engine_info(hfv_ctx* c, hfv_engine_info* out) {
  TRY(require_mesh(c));
  if (!out) return set_err(HFV_ERR_INVALID, "out is NULL");
  // readable after a hard error too (error_code then holds it)
  HIP_OK(hipMemcpyAsync(c->heng, c->deng, sizeof(SimEngine), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  fill_engine_info(c->heng[0], out);
  if (!out->error_code) out->error_code = c->hcnt->err;
  return HFV_OK;
}

int hfv_conservation(hfv_ctx* c, double* out5) {
  TRY(require_mesh(c));
  if (!out5) return set_err(HFV_ERR_INVALID, "out is NULL");
  double* d = nullptr;
  HIP_OK(hipMalloc((void**)&d, sizeof(double) * 5));
  hipLaunchKernelGGL(k_conservation, dim3(1), dim3(1024), 0, c->stream, c->batch_args(), c->cbase[0],
                     c->cbase[0] + c->nused[0], d);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out5, d, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  (void)hipFree(d);
  return HFV_OK;
}

// ---- batches
int hfv_batch_create(int nslots, const int64_t* cell_cap, const int64_t* face_cap, const int64_t* conn_cap,
                     hfv_ctx** out) {
  if (nslots <= 0 || !cell_cap || !face_cap || !conn_cap || !out) return set_err(HFV_ERR_INVALID, "bad batch arguments");
  for (int s = 0; s < nslots; ++s)
    if (cell_cap[s] < 1 || face_cap[s] < 0 || conn_cap[s] < 0)
      return set_err(HFV_ERR_INVALID, "slot capacities: cells >= 1, faces and CSR entries >= 0");
  hfv_ctx* c = nullptr;
  TRY(new_ctx(&c, nslots));
  c->batch = true;
  c->ccap.assign(cell_cap, cell_cap + nslots);
  c->fcap.assign(face_cap, face_cap + nslots);
  c->kcap.assign(conn_cap, conn_cap + nslots);
  int e = alloc_batch(c);
  if (e) {
    hfv_destroy(c);
    return e;
  }
  *out = c;
  return HFV_OK;
}

int hfv_batch_load(hfv_ctx* c, int s, const hfv_params* params, const hfv_mesh_arrays* mesh, const double* loc,
                   const double* mass, const double* vel, const double* u_int) {
  TRY(check_slot(c, s));
  TRY(check_params(params));
  if (!c->batch) return set_err(HFV_ERR_INVALID, "not a batch context");
  if (!mesh) return set_err(HFV_ERR_INVALID, "mesh is NULL");
  for (int k = 0; k < c->nslots; ++k)
    if (k != s && c->loaded[k] && c->params[k].precision != params->precision)
      return set_err(HFV_ERR_INVALID, "all simulations of a batch share one precision");
  HIP_OK(hipStreamSynchronize(c->stream));
  c->params[s] = *params;
  TRY(upload_mesh(c, s, mesh, nullptr, false));
  TRY(reset_slot(c, s));
  c->loaded[s] = 1;
  TRY(write_jobs(c));
  const LaunchSet prev = launch_set(c);
  TRY(init_ic(c, s, mesh->n_cells, loc, mass, vel, u_int));
  recompute_uniforms(c);
  if (!(launch_set(c) == prev)) drop_graph(c);
  return HFV_OK;
}

int hfv_batch_unload(hfv_ctx* c, int s) {
  TRY(check_slot(c, s));
  if (!c->batch) return set_err(HFV_ERR_INVALID, "not a batch context");
  HIP_OK(hipStreamSynchronize(c->stream));
  c->loaded[s] = 0;
  c->nused[s] = 0;
  TRY(write_jobs(c));
  HIP_OK(hipMemcpyAsync(c->heng + s, c->deng + s, sizeof(SimEngine), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->heng[s].done = 1;
  c->heng[s].force = 0;
  HIP_OK(hipMemcpyAsync(c->deng + s, c->heng + s, sizeof(SimEngine), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  const LaunchSet prev = launch_set(c);
  recompute_uniforms(c);
  if (!(launch_set(c) == prev)) drop_graph(c);  // kernels of simulations that left are not launched
  return HFV_OK;
}

int hfv_batch_run(hfv_ctx* c, int64_t max_launches, int64_t* launches) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  HIP_OK(hipSetDevice(c->device));
  int64_t done = 0;
  int64_t chunk = 8;
  while (done < max_launches) {
    int64_t k = std::min<int64_t>(chunk, max_launches - done);
    for (int64_t j = 0; j < k; ++j) TRY(launch_step(c));
    done += k;
    TRY(sync_counters(c));
    prof_collect(c);
    bool all_done = true;
    int64_t lc = 0;
    for (int s = 0; s < c->nslots; ++s)
      if (c->loaded[s] && !c->heng[s].done) {
        all_done = false;
        lc += (c->nused[s] + BLOCK - 1) / BLOCK;
      }
    if (all_done) break;
    // drop finished slots from the grids once they hold a tenth of the launched chunks (each new grid
    // size re-captures the step graph)
    if (10 * lc < 9 * c->live_chunks) TRY(write_jobs(c));
    chunk = std::min<int64_t>(chunk * 2, 64);
  }
  if (launches) *launches = done;
  return HFV_OK;
}

int hfv_batch_get_engine_info(hfv_ctx* c, int s, hfv_engine_info* out) {
  TRY(check_slot(c, s));
  if (!out) return set_err(HFV_ERR_INVALID, "out is NULL");
  HIP_OK(hipMemcpyAsync(c->heng + s, c->deng + s, sizeof(SimEngine), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  fill_engine_info(c->heng[s], out);
  return HFV_OK;
}

int hfv_batch_get_state(hfv_ctx* c, int s, hfv_state* st) {
  TRY(check_slot(c, s));
  if (!st) return set_err(HFV_ERR_INVALID, "state is NULL");
  return copy_state(c, s, st, false);
}

int hfv_batch_num_slots(hfv_ctx* c, int* n) {
  if (!c || !n) return set_err(HFV_ERR_INVALID, "NULL argument");
  *n = c->nslots;
  return HFV_OK;
}

// ---- geometry mode
int hfv_set_geometry_mode(hfv_ctx* c, int use_static) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  if (use_static != 0 && use_static != 1) return set_err(HFV_ERR_INVALID, "mode must be 0 or 1");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  c->use_static = use_static != 0;
  for (int s = 0; s < c->nslots; ++s)
    if (c->loaded[s]) TRY(refresh_geometry(c, s));
  return HFV_OK;
}

int hfv_get_geometry_mode(hfv_ctx* c, int slot, int* is_static) {
  TRY(check_slot(c, slot));
  if (!is_static) return set_err(HFV_ERR_INVALID, "NULL argument");
  *is_static = c->loaded[slot] ? c->desc[slot].static_geom : 0;
  return HFV_OK;
}

int hfv_set_tuning(hfv_ctx* c, int slot, int staged, int static_grad, int gstaged) {
  TRY(check_slot(c, slot));
  if (staged < -1 || staged > 1 || static_grad < -1 || static_grad > 1 || gstaged < -1 || gstaged > 1)
    return set_err(HFV_ERR_INVALID, "tuning values must be -1, 0 or 1");
  HIP_OK(hipStreamSynchronize(c->stream));
  c->tune_stage[slot] = staged;
  c->tune_sgrad[slot] = static_grad;
  c->tune_gstage[slot] = gstaged;
  if (c->loaded[slot]) TRY(refresh_geometry(c, slot));
  return HFV_OK;
}

int hfv_set_launch_shape(hfv_ctx* c, int flux_faces_per_thread) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  const int f = flux_faces_per_thread;
  if (f != 1 && f != 2 && f != 4) return set_err(HFV_ERR_INVALID, "faces per thread must be 1, 2 or 4");
  HIP_OK(hipSetDevice(c->device));
  if (f != c->flux_fpt) {
    HIP_OK(hipStreamSynchronize(c->stream));
    c->flux_fpt = f;
    drop_graph(c);  // the captured step holds the flux grid
  }
  return HFV_OK;
}

int hfv_get_launch_shape(hfv_ctx* c, int* flux_faces_per_thread, double* ns_per_update3) {
  if (!c || !flux_faces_per_thread) return set_err(HFV_ERR_INVALID, "NULL argument");
  *flux_faces_per_thread = c->flux_fpt;
  if (ns_per_update3)
    for (int q = 0; q < 3; ++q) ns_per_update3[q] = c->tune_fpt_ns[q];
  return HFV_OK;
}

int hfv_get_tuning(hfv_ctx* c, int slot, int* staged, int* static_grad, int* gstaged) {
  TRY(check_slot(c, slot));
  if (staged) *staged = c->loaded[slot] ? c->desc[slot].staged : 0;
  if (static_grad) *static_grad = c->loaded[slot] ? c->desc[slot].static_grad : 0;
  if (gstaged) *gstaged = c->loaded[slot] ? c->desc[slot].gstaged : 0;
  return HFV_OK;
}

// Size classes of the autotuner: dimension x cell-count bucket.
static const int64_t kTuneEdges[] = {0, 1024, 8192, 65536, INT64_MAX};
static const double kTuneTie = 0.01;

int hfv_autotune(hfv_ctx* c, int launches, int rounds, hfv_tune_class* out, int max_out, int* n_out) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  if (launches < 1 || rounds < 1) return set_err(HFV_ERR_INVALID, "launches and rounds must be >= 1");
  HIP_OK(hipSetDevice(c->device));
  TRY(sync_counters(c));
  // classes of the live static-geometry simulations
  struct Cls {
    int dim, bucket;
    std::vector<int> slots;
    bool can_stage = false, can_gstage = false;
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int best = -1;
  };
  std::vector<Cls> cls;
  for (int s = 0; s < c->nslots; ++s) {
    if (!c->loaded[s] || c->heng[s].done || !c->desc[s].static_geom) continue;
    int b = 0;
    while (c->nused[s] >= kTuneEdges[b + 1]) ++b;
    const int dim = c->params[s].dim;
    size_t k = 0;
    while (k < cls.size() && !(cls[k].dim == dim && cls[k].bucket == b)) ++k;
    if (k == cls.size()) {
      cls.push_back(Cls{});
      cls[k].dim = dim;
      cls[k].bucket = b;
    }
    cls[k].slots.push_back(s);
    if (c->has_stage[s]) cls[k].can_stage = true;
    // staged gradient neighbours are a 3D candidate only: in 1D and 2D each neighbour line is read
    // 2-6 times and L2 already holds it (timed windows of the ensemble: 0.5-1.9% slower per class)
    if (c->has_gstage[s] && dim == 3) cls[k].can_gstage = true;
  }
  auto apply = [&](const Cls& k, int cand) -> int {
    for (int s : k.slots) {
      c->tune_sgrad[s] = cand & 1;
      c->tune_stage[s] = (cand >> 1) & 1;
      c->tune_gstage[s] = (cand >> 2) & 1;
      TRY(refresh_geometry(c, s));
    }
    return HFV_OK;
  };
  auto updates = [&]() -> long long {
    long long u = 0;
    for (int s = 0; s < c->nslots; ++s)
      if (c->loaded[s]) u += c->heng[s].cell_updates;
    return u;
  };
  auto live_cells = [&]() -> long long {
    long long n = 0;
    for (int s = 0; s < c->nslots; ++s)
      if (c->loaded[s] && !c->heng[s].done) n += c->nused[s];
    return n;
  };
  auto finished = [&]() -> bool { return live_cells() == 0; };
  auto class_done = [&](const Cls& k) -> bool {
    for (int s : k.slots)
      if (!c->heng[s].done) return false;
    return true;
  };
  const bool prof = c->prof;
  c->prof = false;  // graph replays, as in a run
  int rc = HFV_OK;
  bool all_done = false;
  for (auto& k : cls) {
    std::vector<int> cands;
    for (int cand = 0; cand < 8; ++cand)
      if ((k.can_stage || !(cand & 2)) && (k.can_gstage || !(cand & 4))) cands.push_back(cand);
    for (int r = 0; r < rounds && !all_done && rc == HFV_OK; ++r)
      for (int cand : cands) {
        // a class whose simulations have all finished keeps what it has measured (or the default)
        if (class_done(k)) break;
        if ((rc = apply(k, cand)) != HFV_OK) break;
        // one untimed launch: a changed kernel set is captured into a new graph here
        pack_fr_if_dirty(c);
        if ((rc = launch_step(c)) != HFV_OK) break;
        if ((rc = sync_counters(c)) != HFV_OK) break;
        if ((all_done = finished())) break;
        // groups of `launches` launches until the sample holds at least `launches` sweeps of the live
        // cells (at most 4 groups): a run whose steps alternate between many and few active cells
        // would otherwise be measured on whichever steps its launches happened to hit
        const long long u0 = updates(), want = (long long)launches * live_cells();
        double dt = 0.;
        for (int g = 0; g < 4; ++g) {
          const auto t0 = std::chrono::steady_clock::now();
          for (int j = 0; j < launches && rc == HFV_OK; ++j) rc = launch_step(c);
          if (rc != HFV_OK) break;
          // a failed synchronisation ends the tuning like any other error: the restore below still runs
          const hipError_t se = hipStreamSynchronize(c->stream);
          if (se != hipSuccess) {
            rc = set_err(HFV_ERR_HIP, hipGetErrorString(se));
            break;
          }
          dt += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
          if ((rc = sync_counters(c)) != HFV_OK) break;
          if (updates() - u0 >= want || finished()) break;
        }
        if (rc != HFV_OK) break;
        k.t[cand] += dt;
        k.u[cand] += (double)(updates() - u0);
        if ((all_done = finished())) break;
      }
    if (rc != HFV_OK) break;
    double best = 0.;
    for (int cand : cands)
      if (k.u[cand] > 0. && (k.best < 0 || k.t[cand] / k.u[cand] < best)) {
        k.best = cand;
        best = k.t[cand] / k.u[cand];
      }
    // the default stays unless beaten by more than the measurement's own spread
    const int def = (k.can_stage && k.dim != 1 ? 2 : 0) | (k.dim == 3 ? 1 : 0) | (k.dim == 3 && k.can_gstage ? 4 : 0);
    if (k.best < 0 || (k.u[def] > 0. && k.t[def] / k.u[def] <= best * (1. + kTuneTie))) k.best = def;
    if ((rc = apply(k, k.best)) != HFV_OK) break;
  }
  // the flux kernel's launch shape (face chunks per block: faces per thread) for the whole context, with
  // the classes' choices in place; 1 stays unless beaten by more than the tie margin
  const int fpt0 = c->flux_fpt;
  double ft[3] = {0., 0., 0.}, fu[3] = {0., 0., 0.};
  for (int r = 0; r < rounds && !all_done && rc == HFV_OK && !cls.empty(); ++r)
    for (int q = 0; q < 3; ++q) {
      const int fpt = 1 << q;
      if (fpt != c->flux_fpt) {
        c->flux_fpt = fpt;
        drop_graph(c);
      }
      if ((rc = launch_step(c)) != HFV_OK) break;  // untimed: captures the new grid
      if ((rc = sync_counters(c)) != HFV_OK) break;
      if ((all_done = finished())) break;
      const long long u0 = updates(), want = (long long)launches * live_cells();
      double dt = 0.;
      for (int g = 0; g < 4; ++g) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int j = 0; j < launches && rc == HFV_OK; ++j) rc = launch_step(c);
        if (rc != HFV_OK) break;
        const hipError_t se = hipStreamSynchronize(c->stream);
        if (se != hipSuccess) {
          rc = set_err(HFV_ERR_HIP, hipGetErrorString(se));
          break;
        }
        dt += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = sync_counters(c)) != HFV_OK) break;
        if (updates() - u0 >= want || finished()) break;
      }
      if (rc != HFV_OK) break;
      ft[q] += dt;
      fu[q] += (double)(updates() - u0);
      if ((all_done = finished())) break;
    }
  for (int q = 0; q < 3; ++q) c->tune_fpt_ns[q] = fu[q] > 0. ? 1e9 * ft[q] / fu[q] : -1.;
  {
    int bq = 0;
    if (rc == HFV_OK && fu[0] > 0.)
      for (int q = 1; q < 3; ++q)
        if (fu[q] > 0. && ft[q] / fu[q] * (1. + kTuneTie) < ft[bq] / fu[bq]) bq = q;
    const int keep = rc == HFV_OK && fu[0] > 0. ? 1 << bq : fpt0;
    if (keep != c->flux_fpt) {
      c->flux_fpt = keep;
      drop_graph(c);
    }
  }
  c->prof = prof;
  if (rc != HFV_OK) {
    // leave every class on its default choices rather than on the candidate being measured
    for (auto& k : cls)
      for (int s : k.slots) {
        c->tune_sgrad[s] = c->tune_stage[s] = c->tune_gstage[s] = -1;
        (void)refresh_geometry(c, s);
      }
    return rc;
  }
  const int n = (int)cls.size();
  if (n_out) *n_out = n;
  for (int k = 0; k < n && k < max_out && out; ++k) {
    hfv_tune_class& o = out[k];
    o.dim = cls[k].dim;
    o.cells_lo = kTuneEdges[cls[k].bucket];
    o.cells_hi = kTuneEdges[cls[k].bucket + 1];
    o.n_sims = (int)cls[k].slots.size();
    o.staged = (cls[k].best >> 1) & 1;
    o.static_grad = cls[k].best & 1;
    o.gstaged = (cls[k].best >> 2) & 1;
    for (int q = 0; q < 8; ++q) o.ns_per_update[q] = cls[k].u[q] > 0. ? 1e9 * cls[k].t[q] / cls[k].u[q] : -1.;
  }
  return HFV_OK;
}

// ---- measurement
int hfv_profile_enable(hfv_ctx* c, int enable) {
  if (!c) return set_err(HFV_ERR_INVALID, "ctx is NULL");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (enable) {
    for (int r = 0; r < hfv_ctx::PROF_RING; ++r)
      for (int k = 0; k < HFV_NUM_KERNEL_SLOTS; ++k) {
        if (!c->ev0[r][k]) HIP_OK(hipEventCreate(&c->ev0[r][k]));
        if (!c->ev1[r][k]) HIP_OK(hipEventCreate(&c->ev1[r][k]));
        c->pending[r][k] = false;
      }
    c->ring = 0;
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
  TRY(check_params(params));
  if (nb <= 0 || !WL || !WR || !vface || !nrm || !Fo) return set_err(HFV_ERR_INVALID, "bad batch arrays");
  if (nb >= (1ll << 29)) return set_err(HFV_ERR_INVALID, "batch too large for 32-bit indices");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(HFV_ERR_NO_DEVICE, "no HIP device");
  Cfg cfg = make_cfg(*params);
  double *dl, *dr, *dv, *dn, *df;
  DevCounters* cnt;
  HIP_OK(hipMalloc((void**)&dl, sizeof(double) * 5 * nb));
  HIP_OK(hipMalloc((void**)&dr, sizeof(double) * 5 * nb));
  HIP_OK(hipMalloc((void**)&dv, sizeof(double) * 3 * nb));
  HIP_OK(hipMalloc((void**)&dn, sizeof(double) * 3 * nb));
  HIP_OK(hipMalloc((void**)&df, sizeof(double) * 5 * nb));
  HIP_OK(hipMalloc((void**)&cnt, sizeof(DevCounters)));
  DevCounters h;
  std::memset(&h, 0, sizeof(h));
  HIP_OK(hipMemcpy(dl, WL, sizeof(double) * 5 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dr, WR, sizeof(double) * 5 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dv, vface, sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dn, nrm, sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(cnt, &h, sizeof(DevCounters), hipMemcpyHostToDevice));
  if (params->precision == HFV_FP32_FLUX)
    hipLaunchKernelGGL(k_riemann_batch<float>, dim3(grid_full(nb)), dim3(BLOCK), 0, 0, cfg, nb, dl, dr, dv, dn, df, cnt);
  else
    hipLaunchKernelGGL(k_riemann_batch<double>, dim3(grid_full(nb)), dim3(BLOCK), 0, 0, cfg, nb, dl, dr, dv, dn, df, cnt);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(Fo, df, sizeof(double) * 5 * nb, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&h, cnt, sizeof(DevCounters), hipMemcpyDeviceToHost));
  (void)hipFree(dl);
  (void)hipFree(dr);
  (void)hipFree(dv);
  (void)hipFree(dn);
  (void)hipFree(df);
  (void)hipFree(cnt);
  if (h.err) return set_err(h.err, "device raised a hard error in the Riemann solver");
  return HFV_OK;
}

int hfv_riemann_batch_noalloc(const hfv_params* params, int64_t nb, const double* WL, const double* WR,
                              const double* vface, const double* nrm, double* Fo) {
  TRY(check_params(params));
  if (nb <= 0 || !WL || !WR || !vface || !nrm || !Fo) return set_err(HFV_ERR_INVALID, "bad batch arrays");
  if (nb > HFV_NOALLOC_MAX_FACES)
    return set_err(HFV_ERR_INVALID, "hfv_riemann_batch_noalloc: n exceeds HFV_NOALLOC_MAX_FACES (65536 problems per call)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(HFV_ERR_NO_DEVICE, "no HIP device");
  double* base = nullptr;
  TRY(scratch_of_current_device(&base));
  DevCounters* cnt = (DevCounters*)base;
  double* dl = base + kScratchHead;
  double* dr = dl + 5 * nb;
  double* dv = dr + 5 * nb;
  double* dn = dv + 3 * nb;
  double* df = dn + 3 * nb;
  const Cfg cfg = make_cfg(*params);
  DevCounters h;
  std::memset(&h, 0, sizeof(h));
  // the scratch keeps the previous call's contents: the counters start at zero, every staged array is
  // overwritten for [0, n)
  HIP_OK(hipMemcpy(cnt, &h, sizeof(DevCounters), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dl, WL, sizeof(double) * 5 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dr, WR, sizeof(double) * 5 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dv, vface, sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dn, nrm, sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  if (params->precision == HFV_FP32_FLUX)
    hipLaunchKernelGGL(k_riemann_batch<float>, dim3(grid_full(nb)), dim3(BLOCK), 0, 0, cfg, nb, dl, dr, dv, dn, df, cnt);
  else
    hipLaunchKernelGGL(k_riemann_batch<double>, dim3(grid_full(nb)), dim3(BLOCK), 0, 0, cfg, nb, dl, dr, dv, dn, df, cnt);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(Fo, df, sizeof(double) * 5 * nb, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&h, cnt, sizeof(DevCounters), hipMemcpyDeviceToHost));
  if (h.err) return set_err(h.err, "device raised a hard error in the Riemann solver");
  return HFV_OK;
}

}  // extern "C"
