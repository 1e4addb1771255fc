// 2D moving meshes rebuilt on the device: Space::volume_calculation (src/space.rs:236-263) for 2D, the
// reference's call of meshless_voronoi's Voronoi::build_partial.  Included by hfv_api.hip.
//
// The tessellation is the host tessellator's (hydro_rs_amd/csrc/host/voronoi.cpp, which stands in for
// the crate) evaluated with the same operations in the same order, so the device mesh equals the host's
// bit for bit: every generator's cell is the initial region (the box, or a square of K bin widths around
// the generator on a periodic box) clipped by the bisectors of the generators of the (2K+1)^2 bins around
// its own, visited in increasing (squared distance, index, image code) order until no unvisited one can
// cut the cell (a quarter of its squared distance beyond every vertex), with K = 2 .. 6 until the cell
// lies within K bin widths (and, periodic, has no wall edge).  Faces are emitted by the lower-index cell
// of each pair in the order of its polygon's edges; face indices follow the cells; each cell's CSR lists
// its faces in ascending face index.
//
// Per step (simulations with moving2d, while stepping), after the drift and the predict:
//   k_v2_clear, k_v2_bin, k_v2_scan_bins, k_v2_place   the bin grid (counting sort of the generators)
//   k_v2_cells                                         every cell's polygon; the active cells' volume,
//                                                      centroid and search radius (Particle::
//                                                      update_geometry, part.rs:151-158); the cell's
//                                                      emitted faces in a provisional order
//   k_v2_scan_faces, k_v2_faces                        face indices in cell order, the face arrays
//   k_v2_scan_csr, k_v2_csr_fill, k_v2_csr_sort        offsets and CSR entries in ascending face order
// Inactive cells keep their geometry and get new CSR ranges (the reference's note at space.rs:233-235).
#pragma once

namespace hfv {

#define HFV_V2_MAXL 40      // edges of a cell polygon while it is clipped
#define HFV_V2_MAXC 64      // CSR entries of one cell
#define HFV_SCAN_BLOCK 1024

// A face in the provisional order: length, centroid and unit normal relative to the emitting generator,
// the neighbour's local index (< 0: a wall) and periodic image code.
struct V2Face {
  double area, cx, cy, nx, ny;
  int32_t nb, code;
};

struct V2Bufs {
  int32_t* gcnt;   // [N + S + 1], simulation s at cell_base + s: bin counts, then bin starts (nbins + 1)
  int32_t* gfill;  // same layout: placement cursors
  int32_t* gord;   // [N] at cell_base: local cell indices grouped by bin
  int32_t* gkey;   // [N] bin of each cell
  int32_t* femit;  // [N + S + 1]: faces a cell emits, then their exclusive offsets (the total at [n])
  int32_t* vst;    // [N] start of a cell's faces in vf
  int32_t* ccnt;   // [N + S + 1]: CSR entries per cell, then exclusive offsets
  int32_t* cfill;  // [N] CSR fill cursors (global entry index)
  V2Face* vf;      // [F] at face_base: faces in provisional order
};

__device__ __forceinline__ int v2_bin(double x, double w, int nb) {
  int b = (int)floor(x / w);
  if (b < 0) b = 0;
  if (b >= nb) b = nb - 1;
  return b;
}

// (squared distance, index, image code) order; the code compares as the host's memcmp of int8 codes
__device__ __forceinline__ bool v2_less(double ad2, int32_t aj, int32_t ak, double bd2, int32_t bj, int32_t bk) {
  if (ad2 != bd2) return ad2 < bd2;
  if (aj != bj) return aj < bj;
  return ak < bk;
}

__device__ __forceinline__ int32_t v2_shift_code(int s0, int s1) {
  return (s0 == 1 ? 1 : s0 == -1 ? 2 : 0) | ((s1 == 1 ? 1 : s1 == -1 ? 2 : 0) << 2);
}

struct V2Poly {
  double nx[HFV_V2_MAXL], ny[HFV_V2_MAXL], d[HFV_V2_MAXL];
  int32_t nb[HFV_V2_MAXL], code[HFV_V2_MAXL];
  double vx[HFV_V2_MAXL], vy[HFV_V2_MAXL];
  int m;

  // vertex k: the intersection of edges k and k+1 (0 where they are parallel)
  __device__ void recompute() {
    for (int k = 0; k < m; ++k) {
      const int k1 = k + 1 == m ? 0 : k + 1;
      const double det = nx[k] * ny[k1] - ny[k] * nx[k1];
      if (det == 0.0) {
        vx[k] = 0.;
        vy[k] = 0.;
      } else {
        vx[k] = (d[k] * ny[k1] - d[k1] * ny[k]) / det;
        vy[k] = (nx[k] * d[k1] - nx[k1] * d[k]) / det;
      }
    }
  }
  __device__ double rmax2() const {
    double r = 0.;
    for (int k = 0; k < m; ++k) {
      const double q = vx[k] * vx[k] + vy[k] * vy[k];
      r = r < q ? q : r;
    }
    return r;
  }
  __device__ void set(int k, double a, double b, double dd, int32_t n_, int32_t c_) {
    nx[k] = a;
    ny[k] = b;
    d[k] = dd;
    nb[k] = n_;
    code[k] = c_;
  }
  // Clip by the half-plane n.y <= dd: 0 no cut, 1 cut, -1 too many edges.  The run of outside vertices
  // start .. start+nout-1 is contiguous; edges start+1 .. start+nout-1 go, the new edge follows edge
  // start, and the list restarts at the first kept edge after the removed run.
  __device__ int clip(double a, double b, double dd, int32_t n_, int32_t c_, V2Poly& t) {
    uint64_t out = 0;
    int nout = 0;
    for (int k = 0; k < m; ++k)
      if ((a * vx[k] + b * vy[k]) > dd) {
        out |= 1ull << k;
        ++nout;
      }
    if (nout == 0 || nout == m) return 0;
    if (m - nout + 2 > HFV_V2_MAXL) return -1;
    int start = 0;
    for (int k = 0; k < m; ++k) {
      const int km = k == 0 ? m - 1 : k - 1;
      if (((out >> k) & 1) && !((out >> km) & 1)) {
        start = k;
        break;
      }
    }
    int k = (start + nout) % m;
    int q = 0;
    while (true) {
      t.set(q++, nx[k], ny[k], d[k], nb[k], code[k]);
      if (k == start) break;
      k = k + 1 == m ? 0 : k + 1;
    }
    t.set(q++, a, b, dd, n_, c_);
    for (int j = 0; j < q; ++j) set(j, t.nx[j], t.ny[j], t.d[j], t.nb[j], t.code[j]);
    m = q;
    recompute();
    return 1;
  }
};

// ---------------------------------------------------------------- bin grid
__global__ void __launch_bounds__(HFV_BLOCK) k_v2_clear(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.moving2d || !B.eng[s].stepping) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int64_t k = D.cell_base + s + (gi - D.cell_base);
  V.gcnt[k] = 0;
  V.ccnt[k] = 0;
  V.femit[k] = 0;
  if (gi == D.cell_base) {  // entry n (bins: nbins <= n)
    V.gcnt[D.cell_base + s + D.n] = 0;
    V.ccnt[D.cell_base + s + D.n] = 0;
    V.femit[D.cell_base + s + D.n] = 0;
  }
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_bin(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.moving2d || !B.eng[s].stepping) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int64_t N = B.C.n;
  const int key = v2_bin(B.C.loc[gi], D.v2w[0], D.v2nb[0]) * D.v2nb[1] + v2_bin(B.C.loc[N + gi], D.v2w[1], D.v2nb[1]);
  V.gkey[gi] = key;
  atomicAdd(&V.gcnt[D.cell_base + s + key], 1);
}

// Exclusive prefix sum of a[0, len) in place by one block (wave scans, then the waves' totals); returns
// the total to every thread.
__device__ int32_t block_scan_excl(int32_t* a, int64_t len) {
  __shared__ int32_t wsum[HFV_SCAN_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int32_t carry = 0;
  for (int64_t base = 0; base < len; base += blockDim.x) {
    const int64_t k = base + threadIdx.x;
    const int32_t v = k < len ? a[k] : 0;
    int32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
      int32_t w = lane < nw ? wsum[lane] : 0;
      for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(w, o, 64);
        if (lane >= o) w += y;
      }
      if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    const int32_t before = wid > 0 ? wsum[wid - 1] : 0;
    if (k < len) a[k] = carry + before + x - v;
    carry += wsum[nw - 1];
    __syncthreads();
  }
  return carry;
}

// one block per simulation
__global__ void __launch_bounds__(HFV_SCAN_BLOCK) k_v2_scan_bins(Batch B, V2Bufs V) {
  const int s = blockIdx.x;
  const SimDesc& D = B.sims[s];
  if (!D.moving2d || !B.eng[s].stepping) return;
  const int64_t nbins = (int64_t)D.v2nb[0] * D.v2nb[1];
  int32_t* g = V.gcnt + D.cell_base + s;
  block_scan_excl(g, nbins + 1);
  for (int64_t k = threadIdx.x; k < nbins; k += blockDim.x) V.gfill[D.cell_base + s + k] = g[k];
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_place(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.moving2d || !B.eng[s].stepping) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int32_t p = atomicAdd(&V.gfill[D.cell_base + s + V.gkey[gi]], 1);
  V.gord[D.cell_base + p] = (int32_t)(gi - D.cell_base);
}

// ---------------------------------------------------------------- cells
// The cell of generator i with candidate bins within K: 0 built, 1 retry with a larger K, -1 overflow.
__device__ int v2_build_cell(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i, double xi, double yi,
                             double zi, int K, V2Poly& P, V2Poly& T, double& r2) {
  const int64_t N = B.C.n, cb = D.cell_base;
  const double* L = B.C.loc;
  const double bx = D.cfg.box[0], by = D.cfg.box[1], bz = D.cfg.box[2];
  const int nb0 = D.v2nb[0], nb1 = D.v2nb[1];
  const double w0 = D.v2w[0], w1 = D.v2w[1];
  const bool per = D.periodic != 0;
  const int bi0 = v2_bin(xi, w0, nb0), bi1 = v2_bin(yi, w1, nb1);
  const int32_t* gst = V.gcnt + cb + s;
  const int32_t* gord = V.gord + cb;
  // initial region: the box, or K bin widths around the generator on a periodic box
  double reach = 0.;
  reach = reach < K * w0 ? K * w0 : reach;
  reach = reach < K * w1 ? K * w1 : reach;
  const double lox = per ? -reach : -xi, loy = per ? -reach : -yi;
  const double hix = per ? reach : bx - xi, hiy = per ? reach : by - yi;
  P.m = 4;  // counter-clockwise: bottom, right, top, left (walls -3, -2, -4, -1)
  P.set(0, 0., -1., -loy, -3, 0);
  P.set(1, 1., 0., hix, -2, 0);
  P.set(2, 0., 1., hiy, -4, 0);
  P.set(3, -1., 0., -lox, -1, 0);
  P.recompute();
  r2 = P.rmax2();
  bool have_last = false;
  double ld2 = 0.;
  int32_t lj = 0, lk = 0;
  while (true) {
    bool have = false;
    double bd2 = 0., brx = 0., bry = 0.;
    int32_t bj = 0, bk = 0, bcode = 0;
    for (int dx = -K; dx <= K; ++dx) {
      int b0 = bi0 + dx, s0 = 0;
      if (b0 < 0 || b0 >= nb0) {
        if (!per) continue;
        if (b0 < 0) {
          b0 += nb0;
          s0 = -1;
        } else {
          b0 -= nb0;
          s0 = 1;
        }
        if (b0 < 0 || b0 >= nb0) continue;
      }
      for (int dy = -K; dy <= K; ++dy) {
        int b1 = bi1 + dy, s1 = 0;
        if (b1 < 0 || b1 >= nb1) {
          if (!per) continue;
          if (b1 < 0) {
            b1 += nb1;
            s1 = -1;
          } else {
            b1 -= nb1;
            s1 = 1;
          }
          if (b1 < 0 || b1 >= nb1) continue;
        }
        const int key = b0 * nb1 + b1;
        const int32_t kk = ((s0 & 0xFF) << 16) | ((s1 & 0xFF) << 8);
        for (int32_t q = gst[key]; q < gst[key + 1]; ++q) {
          const int32_t j = gord[q];
          if (j == i && s0 == 0 && s1 == 0) continue;
          const int64_t gj = cb + j;
          const double xj = L[gj] + (double)s0 * bx, yj = L[N + gj] + (double)s1 * by, zj = L[2 * N + gj] + 0.0 * bz;
          const double rx = xj - xi, ry = yj - yi, rz = zj - zi;
          const double d2 = rx * rx + ry * ry + rz * rz;
          if (have_last && !v2_less(ld2, lj, lk, d2, j, kk)) continue;
          if (!have || v2_less(d2, j, kk, bd2, bj, bk)) {
            have = true;
            bd2 = d2;
            bj = j;
            bk = kk;
            brx = rx;
            bry = ry;
            bcode = v2_shift_code(s0, s1);
          }
        }
      }
    }
    if (!have || 0.25 * bd2 > r2) break;  // security radius: the bisector lies beyond every vertex
    const int c = P.clip(brx, bry, 0.5 * bd2, bj, bcode, T);
    if (c < 0) return -1;
    if (c > 0) r2 = P.rmax2();
    have_last = true;
    ld2 = bd2;
    lj = bj;
    lk = bk;
  }
  // completeness: every unvisited generator lies farther than K bin widths
  double minw = 1e300;
  minw = K * w0 < minw ? K * w0 : minw;
  minw = K * w1 < minw ? K * w1 : minw;
  if (4.0 * r2 > minw * minw) return 1;
  if (per) {  // a periodic cell must not touch the initial region
    for (int k = 0; k < P.m; ++k) {
      const int km = k == 0 ? P.m - 1 : k - 1;
      const double ex = P.vx[k] - P.vx[km], ey = P.vy[k] - P.vy[km];
      if (P.nb[k] < 0 && sqrt(ex * ex + ey * ey) > 0.) return 1;
    }
  }
  return 0;
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_cells(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (!D.moving2d || !E.stepping) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int64_t N = B.C.n, cb = D.cell_base;
  const int32_t i = (int32_t)(gi - cb);
  const double xi = B.C.loc[gi], yi = B.C.loc[N + gi], zi = B.C.loc[2 * N + gi];
  V2Poly P, T;
  double r2 = 0.;
  int r = 1;
  for (int K = 2; K <= 6 && r == 1; ++K) r = v2_build_cell(B, D, V, s, i, xi, yi, zi, K, P, T, r2);
  if (r != 0) {
    V.femit[cb + s + i] = 0;
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
    return;
  }
  // area, centroid and edges (the host tessellator's Cell2::finish)
  const int m = P.m;
  double area = 0., cx = 0., cy = 0.;
  int cnt = 0;
  for (int k = 0; k < m; ++k) {
    const int km = k == 0 ? m - 1 : k - 1;
    const double ax = P.vx[km], ay = P.vy[km], bxx = P.vx[k], byy = P.vy[k];
    const double cr = ax * byy - bxx * ay;
    area += 0.5 * cr;
    cx = cx + (cr / 6.0) * (ax + bxx);
    cy = cy + (cr / 6.0) * (ay + byy);
    const double len = sqrt((bxx - ax) * (bxx - ax) + (byy - ay) * (byy - ay));
    if (len > 0. && (P.nb[k] < 0 || P.nb[k] > i)) ++cnt;
  }
  const int32_t start = atomicAdd(&E.nf_tmp, cnt);
  if ((int64_t)start + cnt > D.face_cap) {
    V.femit[cb + s + i] = 0;
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
    return;
  }
  V.vst[gi] = start;
  V.femit[cb + s + i] = cnt;
  V2Face* out = V.vf + D.face_base + start;
  int q = 0;
  for (int k = 0; k < m; ++k) {
    const int km = k == 0 ? m - 1 : k - 1;
    const double ax = P.vx[km], ay = P.vy[km], bxx = P.vx[k], byy = P.vy[k];
    const double len = sqrt((bxx - ax) * (bxx - ax) + (byy - ay) * (byy - ay));
    if (!(len > 0.) || !(P.nb[k] < 0 || P.nb[k] > i)) continue;
    const double inv = 1.0 / sqrt(P.nx[k] * P.nx[k] + P.ny[k] * P.ny[k]);
    V2Face f;
    f.area = len;
    f.cx = 0.5 * (ax + bxx);
    f.cy = 0.5 * (ay + byy);
    f.nx = inv * P.nx[k];
    f.ny = inv * P.ny[k];
    f.nb = P.nb[k] < 0 ? -1 : P.nb[k];
    f.code = P.code[k];
    out[q++] = f;
  }
  if (B.C.act[gi]) {  // Particle::update_geometry (part.rs:151-158)
    const double rmax = sqrt(P.rmax2());
    B.C.vol[gi] = area;
    const double ia = area > 0. ? 1.0 / area : 0.;
    B.C.cen[gi] = xi + (area > 0. ? ia * cx : 0.);
    B.C.cen[N + gi] = yi + (area > 0. ? ia * cy : 0.);
    B.C.cen[2 * N + gi] = zi;
    B.C.srad[gi] = 1.5 * 2.0 * rmax;
  }
}

// ---------------------------------------------------------------- faces and CSR
__global__ void __launch_bounds__(HFV_SCAN_BLOCK) k_v2_scan_faces(Batch B, V2Bufs V) {
  const int s = blockIdx.x;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (!D.moving2d || !E.stepping) return;
  const int32_t tot = block_scan_excl(V.femit + D.cell_base + s, D.n + 1);
  if (threadIdx.x == 0) {
    // a failed cell emitted nothing: the mesh is incomplete, the simulation stops with E_MESH
    E.nf_mesh = E.err ? 0 : tot;
  }
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_faces(Batch B, V2Bufs V) {
  const BlockJob J = B.fjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  const SimEngine& E = B.eng[s];
  if (!D.moving2d || !E.stepping) return;
  const int64_t f = (int64_t)J.lo + threadIdx.x;
  if (f >= J.hi) return;
  Faces& F = B.F;
  const int64_t fl = f - D.face_base, cb = D.cell_base, NF = F.nf, N = B.C.n;
  if (fl >= E.nf_mesh) {
    F.left[f] = -1;
    F.right[f] = -1;
    return;
  }
  const int32_t* off = V.femit + cb + s;
  int64_t lo = 0, hi = D.n;  // off[lo] <= fl < off[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= fl) lo = mid;
    else hi = mid;
  }
  const int64_t gi = cb + lo;
  const V2Face r = V.vf[D.face_base + V.vst[gi] + (fl - off[lo])];
  F.area[f] = r.area;
  F.nrm[f] = r.nx;
  F.nrm[NF + f] = r.ny;
  F.nrm[2 * NF + f] = 0.;
  F.fcen[f] = B.C.loc[gi] + r.cx;
  F.fcen[NF + f] = B.C.loc[N + gi] + r.cy;
  F.fcen[2 * NF + f] = B.C.loc[2 * N + gi];
  F.left[f] = (int32_t)gi;
  F.right[f] = r.nb < 0 ? -1 : (int32_t)(cb + r.nb);
  F.shc[f] = (uint8_t)r.code;
  atomicAdd(&V.ccnt[cb + s + lo], 1);
  if (r.nb >= 0) atomicAdd(&V.ccnt[cb + s + r.nb], 1);
}

__global__ void __launch_bounds__(HFV_SCAN_BLOCK) k_v2_scan_csr(Batch B, V2Bufs V) {
  const int s = blockIdx.x;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (!D.moving2d || !E.stepping) return;
  const int64_t cb = D.cell_base, n = D.n;
  int32_t* c = V.ccnt + cb + s;
  const int32_t tot = block_scan_excl(c, n + 1);
  if (tot > D.conn_cap) {
    if (threadIdx.x == 0) {
      atomicCAS(&B.cnt->err, 0, E_MESH);
      atomicCAS(&E.err, 0, E_MESH);
    }
    for (int64_t k = threadIdx.x; k < D.noff; k += blockDim.x) B.F.off[cb + k] = (int32_t)D.conn_base;
    return;
  }
  for (int64_t k = threadIdx.x; k < D.noff; k += blockDim.x) {
    const int32_t o = (int32_t)D.conn_base + (k <= n ? c[k] : tot);
    B.F.off[cb + k] = o;
    if (k < n) V.cfill[cb + k] = o;
  }
  if (tot > D.conn_cap) E.nf_mesh = 0;
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_csr_fill(Batch B, V2Bufs V) {
  const BlockJob J = B.fjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  const SimEngine& E = B.eng[s];
  if (!D.moving2d || !E.stepping || E.err) return;
  const int64_t f = (int64_t)J.lo + threadIdx.x;
  if (f >= J.hi || f - D.face_base >= E.nf_mesh) return;
  const int32_t l = B.F.left[f], r = B.F.right[f];
  int32_t p = atomicAdd(&V.cfill[l], 1);
  B.F.conn[p] = (int32_t)(f << 1);
  B.F.nbr[p] = r;
  if (r >= 0) {
    p = atomicAdd(&V.cfill[r], 1);
    B.F.conn[p] = (int32_t)((f << 1) | 1);
    B.F.nbr[p] = l;
  }
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_csr_sort(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (!D.moving2d || !E.stepping || E.err) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int32_t e0 = B.F.off[gi], e1 = B.F.off[gi + 1];
  if (e1 - e0 > HFV_V2_MAXC) {
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
    return;
  }
  int32_t c[HFV_V2_MAXC], nb[HFV_V2_MAXC];
  const int m = e1 - e0;
  for (int k = 0; k < m; ++k) {
    c[k] = B.F.conn[e0 + k];
    nb[k] = B.F.nbr[e0 + k];
  }
  for (int k = 1; k < m; ++k) {  // insertion sort by face index (then side)
    const int32_t x = c[k], y = nb[k];
    int j = k - 1;
    while (j >= 0 && c[j] > x) {
      c[j + 1] = c[j];
      nb[j + 1] = nb[j];
      --j;
    }
    c[j + 1] = x;
    nb[j + 1] = y;
  }
  for (int k = 0; k < m; ++k) {
    B.F.conn[e0 + k] = c[k];
    B.F.nbr[e0 + k] = nb[k];
  }
}

}  // namespace hfv
