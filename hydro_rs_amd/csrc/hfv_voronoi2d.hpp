// 2D moving meshes rebuilt on the device: Space::volume_calculation (src/space.rs:236-263) for 2D, the
// reference's call of meshless_voronoi's Voronoi::build_partial with the step's active mask: only the
// active cells are built (their faces, including those with inactive neighbours), and every cell gets
// its CSR range over the faces that exist.  Included by hfv_api.hip.
//
// The tessellation is the host tessellator's (hydro_rs_amd/csrc/host/voronoi.cpp, which stands in for
// the crate) evaluated with the same operations in the same order, so the device mesh equals the host's
// bit for bit: every generator's cell is the initial region (the box, or a square of K bin widths around
// the generator on a periodic box) clipped by the bisectors of the generators of the (2K+1)^2 bins around
// its own, visited in increasing (squared distance, index, image code) order until no unvisited one can
// cut the cell (a quarter of its squared distance beyond every vertex), with K = 2, 3, .. (up to the whole
// grid) until the cell lies within K bin widths (and, periodic, has no wall edge).  Faces are emitted in
// the order of the emitting cell's polygon edges (v2_emits); face indices follow the cells; each cell's
// CSR lists its faces in ascending face index.
//
// Per step (simulations with vmesh = 2, while stepping; 3D: hfv_voronoi3d.hpp), after the drift and the predict:
//   k_v2_clear, k_v2_bin, scan<0>, k_v2_place          the bin grid (counting sort of the generators)
//   k_v2_cells<0>, k_v2_cells<1>                       the active cells' polygons, volume,
//                                                      centroid and search radius (Particle::
//                                                      update_geometry, part.rs:151-158); the cell's
//                                                      emitted faces in a provisional order
//   scan<1>, k_v2_faces                                face indices in cell order, the face arrays
//   scan<2>, k_v2_csr_fill, k_v2_csr_sort              offsets and CSR entries in ascending face order
// (scan<A>: k_v2_scan_up, k_v2_scan_mid, k_v2_scan_down)
// Inactive cells keep their geometry and get new CSR ranges (the reference's note at space.rs:233-235).
#pragma once

namespace hfv {

#define HFV_SCAN_BLOCK 1024
// A clip removes the vertices beyond a plane by more than this fraction of its offset (the host
// tessellator's threshold): vertices on it test as beyond by a rounding error on exact lattices.
#define HFV_CLIP_TOL 1e-12

// A face in the provisional order: area (length in 2D), centroid and unit normal relative to the emitting
// generator, the neighbour's local index (< 0: a wall) and periodic image code.
struct V2Face {
  double area, cx, cy, cz, nx, ny, nz;
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
  int32_t* part;   // [cell chunks] chunk sums, then chunk offsets, of the prefix sums
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

// ---------------------------------------------------------------- bin grid
__global__ void __launch_bounds__(HFV_BLOCK) k_v2_clear(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.vmesh || !B.eng[s].stepping) return;
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
  if (!D.vmesh || !B.eng[s].stepping) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int64_t N = B.C.n;
  const int key = (v2_bin(B.C.loc[gi], D.vw[0], D.vnb[0]) * D.vnb[1] + v2_bin(B.C.loc[N + gi], D.vw[1], D.vnb[1])) *
                      D.vnb[2] +
                  (D.vmesh == 3 ? v2_bin(B.C.loc[2 * N + gi], D.vw[2], D.vnb[2]) : 0);
  V.gkey[gi] = key;
  atomicAdd(&V.gcnt[D.cell_base + s + key], 1);
}

// ---------------------------------------------------------------- prefix sums
// Exclusive prefix sums over a simulation's region of one of three arrays, in three launches over the
// cell block table (chunks of HFV_BLOCK elements): k_v2_scan_up sums each chunk, k_v2_scan_mid (one block
// per simulation) scans the chunk sums and stores the total after the region, k_v2_scan_down scans each
// chunk from its offset.  A: 0 bin counts (length nbins; the bin starts, copied to the placement
// cursors), 1 emitted faces per cell (length n; the face offsets, the total is the mesh's face count),
// 2 CSR entries per cell (length n; the CSR offsets of the cells).
template <int A>
__device__ __forceinline__ int32_t* v2_scan_region(const V2Bufs& V, const SimDesc& D, int s, int64_t& len) {
  len = A == 0 ? (int64_t)D.vnb[0] * D.vnb[1] * D.vnb[2] : D.n;
  return (A == 0 ? V.gcnt : A == 1 ? V.femit : V.ccnt) + D.cell_base + s;
}

// exclusive scan of one value per thread over a block of HFV_BLOCK threads; *total receives the sum
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* total) {
  __shared__ int32_t ws[HFV_BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[wid] = x;
  __syncthreads();
  int32_t before = 0, tot = 0;
  for (int w = 0; w < HFV_BLOCK / 64; ++w) {
    if (w < wid) before += ws[w];
    tot += ws[w];
  }
  __syncthreads();
  *total = tot;
  return before + x - v;
}

template <int A>
__global__ void __launch_bounds__(HFV_BLOCK) k_v2_scan_up(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.vmesh || !B.eng[s].stepping) return;
  int64_t len;
  const int32_t* a = v2_scan_region<A>(V, D, s, len);
  const int64_t k = (int64_t)J.lo - D.cell_base + threadIdx.x;
  int32_t tot;
  block_excl_scan(k < len ? a[k] : 0, &tot);
  if (threadIdx.x == 0) V.part[J.chunk] = tot;
}

template <int A>
__global__ void __launch_bounds__(HFV_SCAN_BLOCK) k_v2_scan_mid(Batch B, V2Bufs V) {
  const int s = blockIdx.x;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (!D.vmesh || !E.stepping) return;
  int64_t len;
  int32_t* a = v2_scan_region<A>(V, D, s, len);
  __shared__ int32_t wsum[HFV_SCAN_BLOCK / 64];
  __shared__ int32_t tot_s;
  // the chunk sums, HFV_SCAN_BLOCK at a time: wave scans, then the waves' totals
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int32_t carry = 0;
  for (int32_t c0 = D.cchunk0; c0 < D.cchunk1; c0 += blockDim.x) {
    const int32_t c = c0 + (int32_t)threadIdx.x;
    const int32_t v = c < D.cchunk1 ? V.part[c] : 0;
    int32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int32_t before = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
      if (w < wid) before += wsum[w];
      tot += wsum[w];
    }
    if (c < D.cchunk1) V.part[c] = carry + before + x - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a[len] = carry;
    tot_s = carry;
    if (A == 1) E.nf_mesh = E.err ? 0 : carry;  // a failed cell emitted nothing: the simulation stops
    if (A == 2 && carry > D.conn_cap) {
      atomicCAS(&B.cnt->err, 0, E_MESH);
      atomicCAS(&E.err, 0, E_MESH);
    }
  }
  __syncthreads();
  if (A == 2) {  // offsets past the last cell: the slot's end of the CSR
    const int32_t end = (int32_t)D.conn_base + (E.err ? 0 : tot_s);
    for (int64_t k = D.n + threadIdx.x; k < D.noff; k += blockDim.x) B.F.off[D.cell_base + k] = end;
  }
}

template <int A>
__global__ void __launch_bounds__(HFV_BLOCK) k_v2_scan_down(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  const SimEngine& E = B.eng[s];
  if (!D.vmesh || !E.stepping) return;
  int64_t len;
  int32_t* a = v2_scan_region<A>(V, D, s, len);
  const int64_t k = (int64_t)J.lo - D.cell_base + threadIdx.x;
  int32_t tot;
  const int32_t x = V.part[J.chunk] + block_excl_scan(k < len ? a[k] : 0, &tot);
  if (k >= len) return;
  a[k] = x;
  if (A == 0) V.gfill[D.cell_base + s + k] = x;
  if (A == 2) {
    const int32_t o = (int32_t)D.conn_base + (E.err ? 0 : x);
    B.F.off[D.cell_base + k] = o;
    V.cfill[D.cell_base + k] = o;
  }
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_place(Batch B, V2Bufs V) {
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  if (!D.vmesh || !B.eng[s].stepping) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;
  const int32_t p = atomicAdd(&V.gfill[D.cell_base + s + V.gkey[gi]], 1);
  V.gord[D.cell_base + p] = (int32_t)(gi - D.cell_base);
}

// ---------------------------------------------------------------- cells, a team of lanes per cell
// HFV_V2_TEAM lanes of one wavefront build one cell together.  The polygon lives in LDS, each lane
// testing, moving and intersecting every TEAM-th edge; the next candidate in (squared distance, index,
// image code) order is a team minimum over the lanes' smallest keys above the last one, each lane taking
// every TEAM-th bin of the (2K+1)^2 block.  k_v2_cells keeps each lane's candidates in registers (at most
// HFV_V2_TC) and polygons of up to HFV_V2_TL edges; cells that outgrow either go to k_v2_cells_wide,
// which rescans the bins for every candidate and holds polygons of up to HFV_V2_TLW edges.  Both make the
// host tessellator's operations in its order, so the same bits.
// 16 lanes per cell: 205 us per step on Sedov 256^2 with steered generators, against 325 with 8 and 217
// with 32 (the rebuild is latency-bound: few active cells per step, each team's chain of clips)
#ifndef HFV_V2_TEAM
#define HFV_V2_TEAM 16
#endif
#ifndef HFV_V2_TL
#define HFV_V2_TL 32
#endif
#ifndef HFV_V2_TLW
#define HFV_V2_TLW 48
#endif
#ifndef HFV_V2_TC
#define HFV_V2_TC 8
#endif
#define HFV_V2_SUB HFV_V2_TEAM  // blocks per cell chunk
#define HFV_V2_WIDE (-1)         // femit mark: the cell goes to k_v2_cells_wide

template <int TL>
struct V2TeamPoly {
  double nx[TL], ny[TL], d[TL], vx[TL], vy[TL];
  int32_t nb[TL], code[TL];
};

// Voronoi::build_partial builds the active cells only: a face is emitted by the lower-index cell of a pair
// of active cells, by the active cell of a pair with an inactive one, and a wall by its cell.
__device__ __forceinline__ bool v2_emits(const Batch& B, int64_t cb, int32_t nb, int32_t i) {
  return nb < 0 || nb > i || !B.C.act[cb + nb];
}

// The finished cell: area and centroid summed over the edges in order (the host tessellator's
// Cell2::finish), the emitted faces appended to vf in edge order, and the cell's geometry
// (Particle::update_geometry, part.rs:151-158).
template <int TL>
__device__ void v2_finish(const Batch& B, const V2Bufs& V, const SimDesc& D, SimEngine& E, int s, int64_t gi,
                          int32_t i, double xi, double yi, double zi, const V2TeamPoly<TL>& P, int m) {
  const int64_t N = B.C.n, cb = D.cell_base;
  double area = 0., cx = 0., cy = 0., r2 = 0.;
  int cnt = 0;
  for (int k = 0; k < m; ++k) {
    const int km = k == 0 ? m - 1 : k - 1;
    const double ax = P.vx[km], ay = P.vy[km], bxx = P.vx[k], byy = P.vy[k];
    const double cr = ax * byy - bxx * ay;
    area += 0.5 * cr;
    cx = cx + (cr / 6.0) * (ax + bxx);
    cy = cy + (cr / 6.0) * (ay + byy);
    const double len = sqrt((bxx - ax) * (bxx - ax) + (byy - ay) * (byy - ay));
    if (len > 0. && v2_emits(B, cb, P.nb[k], i)) ++cnt;
    const double q = bxx * bxx + byy * byy;
    r2 = r2 < q ? q : r2;
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
    if (!(len > 0.) || !v2_emits(B, cb, P.nb[k], i)) continue;
    const double inv = 1.0 / sqrt(P.nx[k] * P.nx[k] + P.ny[k] * P.ny[k]);
    V2Face f;
    f.area = len;
    f.cx = 0.5 * (ax + bxx);
    f.cy = 0.5 * (ay + byy);
    f.cz = 0.;
    f.nx = inv * P.nx[k];
    f.ny = inv * P.ny[k];
    f.nz = 0.;
    f.nb = P.nb[k] < 0 ? -1 : P.nb[k];
    f.code = P.code[k];
    out[q++] = f;
  }
  if (B.C.act[gi]) {
    B.C.vol[gi] = area;
    const double ia = area > 0. ? 1.0 / area : 0.;
    B.C.cen[gi] = xi + (area > 0. ? ia * cx : 0.);
    B.C.cen[N + gi] = yi + (area > 0. ? ia * cy : 0.);
    B.C.cen[2 * N + gi] = zi;
    B.C.srad[gi] = 1.5 * 2.0 * sqrt(r2);
  }
}

// LDS writes of a team visible to the team's next reads (a team is part of one wavefront)
__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t team_ballot(bool p, int tiw) {
  return (uint32_t)((__ballot(p) >> (HFV_V2_TEAM * tiw)) & ((1ull << HFV_V2_TEAM) - 1));
}

// vertex k: the intersection of edges k and k+1 (0 where they are parallel)
template <int TL>
__device__ __forceinline__ void team_recompute(V2TeamPoly<TL>& P, int m, int lane) {
#pragma unroll
  for (int t = 0; t < TL; t += HFV_V2_TEAM) {
    const int k = t + lane;
    if (k < m) {
      const int k1 = k + 1 == m ? 0 : k + 1;
      const double nx0 = P.nx[k], ny0 = P.ny[k], d0 = P.d[k], nx1 = P.nx[k1], ny1 = P.ny[k1], d1 = P.d[k1];
      const double det = nx0 * ny1 - ny0 * nx1;
      if (det == 0.0) {
        P.vx[k] = 0.;
        P.vy[k] = 0.;
      } else {
        P.vx[k] = (d0 * ny1 - d1 * ny0) / det;
        P.vy[k] = (nx0 * d1 - nx1 * d0) / det;
      }
    }
  }
  team_sync();
}

template <int TL>
__device__ __forceinline__ double team_rmax2(const V2TeamPoly<TL>& P, int m, int lane) {
  double r = 0.;
#pragma unroll
  for (int t = 0; t < TL; t += HFV_V2_TEAM) {
    const int k = t + lane;
    if (k < m) {
      const double q = P.vx[k] * P.vx[k] + P.vy[k] * P.vy[k];
      r = r < q ? q : r;
    }
  }
#pragma unroll
  for (int o = HFV_V2_TEAM / 2; o > 0; o >>= 1) {
    const double q = __shfl_xor(r, o, HFV_V2_TEAM);
    r = r < q ? q : r;
  }
  return r;
}

// Clip by the half-plane n.y <= dd: 0 no cut, 1 cut, -1 more than TL edges.  The run of outside vertices
// start .. start+nout-1 is contiguous; edges start+1 .. start+nout-1 go, the new edge follows edge start,
// and the list restarts at the first kept edge after the removed run.
template <int TL>
__device__ __forceinline__ int team_clip(V2TeamPoly<TL>& P, int& m, double a, double b, double dd, int32_t n_,
                                         int32_t c_, int lane, int tiw) {
  static_assert(TL <= 64 && TL % HFV_V2_TEAM == 0, "edge masks");
  uint64_t om = 0;
#pragma unroll
  for (int t = 0; t < TL; t += HFV_V2_TEAM) {
    const int k = t + lane;
    const bool o = k < m && (a * P.vx[k] + b * P.vy[k]) > dd + HFV_CLIP_TOL * dd;
    om |= (uint64_t)team_ballot(o, tiw) << t;
  }
  const int nout = __popcll(om);
  if (nout == 0 || nout == m) return 0;
  const int q = m - nout + 2;
  if (q > TL) return -1;
  const uint64_t full = m == 64 ? ~0ull : (1ull << m) - 1ull;
  const uint64_t prev = ((om << 1) | (om >> (m - 1))) & full;  // bit k: vertex k-1 is outside
  const uint64_t sm = om & ~prev;
  const int start = sm ? __builtin_ctzll(sm) : 0;
  const int keep = (start + nout) % m;
  constexpr int R = TL / HFV_V2_TEAM;
  double rx[R], ry[R], rd[R];
  int32_t rn[R], rc[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int p = t * HFV_V2_TEAM + lane;
    if (p < q - 1) {
      int o = keep + p;
      o = o >= m ? o - m : o;
      rx[t] = P.nx[o];
      ry[t] = P.ny[o];
      rd[t] = P.d[o];
      rn[t] = P.nb[o];
      rc[t] = P.code[o];
    }
  }
  team_sync();
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int p = t * HFV_V2_TEAM + lane;
    if (p < q - 1) {
      P.nx[p] = rx[t];
      P.ny[p] = ry[t];
      P.d[p] = rd[t];
      P.nb[p] = rn[t];
      P.code[p] = rc[t];
    }
  }
  if (lane == 0) {
    P.nx[q - 1] = a;
    P.ny[q - 1] = b;
    P.d[q - 1] = dd;
    P.nb[q - 1] = n_;
    P.code[q - 1] = c_;
  }
  team_sync();
  m = q;
  team_recompute(P, m, lane);
  return 1;
}

// The candidate key and plane of one generator image (bins wrapped once on a periodic box).
struct V2Cand {
  double d2, rx, ry;
  int32_t j, k, code;
};

// Visit this lane's candidates of the (2K+1)^2 block: f(cand) for each.
template <class Fn>
__device__ __forceinline__ void v2_lane_candidates(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i,
                                                   double xi, double yi, double zi, int K, int lane, Fn&& f) {
  const int64_t N = B.C.n, cb = D.cell_base;
  const double* L = B.C.loc;
  const double bx = D.cfg.box[0], by = D.cfg.box[1], bz = D.cfg.box[2];
  const int nb0 = D.vnb[0], nb1 = D.vnb[1];
  const bool per = D.periodic != 0;
  const int bi0 = v2_bin(xi, D.vw[0], nb0), bi1 = v2_bin(yi, D.vw[1], nb1);
  const int32_t* gst = V.gcnt + cb + s;
  const int32_t* gord = V.gord + cb;
  const int side = 2 * K + 1;
  for (int bb = lane; bb < side * side; bb += HFV_V2_TEAM) {
    int b0 = bi0 + bb / side - K, b1 = bi1 + bb % side - K, s0 = 0, s1 = 0;
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
    const int32_t kk = ((s0 & 0xFF) << 16) | ((s1 & 0xFF) << 8);  // the host's memcmp order of the int8 codes
    const int32_t code = v2_shift_code(s0, s1);
    for (int32_t q = gst[key]; q < gst[key + 1]; ++q) {
      const int32_t j = gord[q];
      if (j == i && s0 == 0 && s1 == 0) continue;
      const int64_t gj = cb + j;
      const double xj = L[gj] + (double)s0 * bx, yj = L[N + gj] + (double)s1 * by, zj = L[2 * N + gj] + 0.0 * bz;
      V2Cand c;
      c.rx = xj - xi;
      c.ry = yj - yi;
      const double rz = zj - zi;
      c.d2 = c.rx * c.rx + c.ry * c.ry + rz * rz;
      c.j = j;
      c.k = kk;
      c.code = code;
      f(c);
    }
  }
}

// Team minimum of the lanes' best candidates (unique keys: every lane ends with the same one).
__device__ __forceinline__ void team_min(bool& have, V2Cand& b) {
#pragma unroll
  for (int o = HFV_V2_TEAM / 2; o > 0; o >>= 1) {
    const int oh = __shfl_xor((int)have, o, HFV_V2_TEAM);
    V2Cand c;
    c.d2 = __shfl_xor(b.d2, o, HFV_V2_TEAM);
    c.rx = __shfl_xor(b.rx, o, HFV_V2_TEAM);
    c.ry = __shfl_xor(b.ry, o, HFV_V2_TEAM);
    c.j = __shfl_xor(b.j, o, HFV_V2_TEAM);
    c.k = __shfl_xor(b.k, o, HFV_V2_TEAM);
    c.code = __shfl_xor(b.code, o, HFV_V2_TEAM);
    if (oh && (!have || v2_less(c.d2, c.j, c.k, b.d2, b.j, b.k))) {
      have = true;
      b = c;
    }
  }
}

// The cell of generator i with candidate bins within K: 0 built, 1 retry with a larger K, 2 too many
// candidates or edges for this variant (LIST: candidates kept in registers; otherwise rescanned).
template <bool LIST, int TL>
__device__ int v2_team_build(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i, double xi,
                             double yi, double zi, int K, V2TeamPoly<TL>& P, int& m, double& r2, int lane, int tiw) {
  const double w0 = D.vw[0], w1 = D.vw[1];
  const bool per = D.periodic != 0;
  // initial region: the box, or K bin widths around the generator on a periodic box
  double reach = 0.;
  reach = reach < K * w0 ? K * w0 : reach;
  reach = reach < K * w1 ? K * w1 : reach;
  const double lox = per ? -reach : -xi, loy = per ? -reach : -yi;
  const double hix = per ? reach : D.cfg.box[0] - xi, hiy = per ? reach : D.cfg.box[1] - yi;
  if (lane < 4) {  // counter-clockwise: bottom, right, top, left (walls -3, -2, -4, -1)
    P.nx[lane] = lane == 1 ? 1. : lane == 3 ? -1. : 0.;
    P.ny[lane] = lane == 0 ? -1. : lane == 2 ? 1. : 0.;
    P.d[lane] = lane == 0 ? -loy : lane == 1 ? hix : lane == 2 ? hiy : -lox;
    P.nb[lane] = lane == 0 ? -3 : lane == 1 ? -2 : lane == 2 ? -4 : -1;
    P.code[lane] = 0;
  }
  team_sync();
  m = 4;
  team_recompute(P, m, lane);
  r2 = team_rmax2(P, m, lane);
  V2Cand cl[LIST ? HFV_V2_TC : 1];
  int cnt = 0;
  if (LIST) {
    v2_lane_candidates(B, D, V, s, i, xi, yi, zi, K, lane, [&](const V2Cand& c) {
#pragma unroll
      for (int q = 0; q < (LIST ? HFV_V2_TC : 1); ++q)
        if (q == cnt) cl[q] = c;
      ++cnt;
    });
    if (team_ballot(cnt > HFV_V2_TC, tiw)) return 2;
  }
  bool have_last = false;
  V2Cand last{};
  while (true) {
    bool have = false;
    V2Cand best{};
    auto consider = [&](const V2Cand& c) {
      if ((!have_last || v2_less(last.d2, last.j, last.k, c.d2, c.j, c.k)) &&
          (!have || v2_less(c.d2, c.j, c.k, best.d2, best.j, best.k))) {
        have = true;
        best = c;
      }
    };
    if (LIST) {
#pragma unroll
      for (int q = 0; q < (LIST ? HFV_V2_TC : 1); ++q)
        if (q < cnt) consider(cl[q]);
    } else {
      v2_lane_candidates(B, D, V, s, i, xi, yi, zi, K, lane, consider);
    }
    team_min(have, best);
    if (!have || 0.25 * best.d2 > r2) break;  // security radius: the bisector lies beyond every vertex
    const int c = team_clip(P, m, best.rx, best.ry, 0.5 * best.d2, best.j, best.code, lane, tiw);
    if (c < 0) return 2;
    if (c > 0) r2 = team_rmax2(P, m, lane);
    have_last = true;
    last = best;
  }
  // completeness: every unvisited generator lies farther than K bin widths
  double minw = 1e300;
  minw = K * w0 < minw ? K * w0 : minw;
  minw = K * w1 < minw ? K * w1 : minw;
  if (4.0 * r2 > minw * minw) return 1;
  if (per) {  // a periodic cell must not touch the initial region
    bool wall = false;
#pragma unroll
    for (int t = 0; t < TL; t += HFV_V2_TEAM) {
      const int k = t + lane;
      if (k < m) {
        const int km = k == 0 ? m - 1 : k - 1;
        const double ex = P.vx[k] - P.vx[km], ey = P.vy[k] - P.vy[km];
        wall = wall || (P.nb[k] < 0 && sqrt(ex * ex + ey * ey) > 0.);
      }
    }
    if (team_ballot(wall, tiw)) return 1;
  }
  return 0;
}

// grid: HFV_V2_SUB blocks per cell chunk, HFV_BLOCK / HFV_V2_TEAM cells per block.  WIDE: only the cells
// that the first pass marked.
template <bool WIDE>
__global__ void __launch_bounds__(HFV_BLOCK) k_v2_cells(Batch B, V2Bufs V) {
  constexpr int TL = WIDE ? HFV_V2_TLW : HFV_V2_TL;
  __shared__ V2TeamPoly<TL> polys[HFV_BLOCK / HFV_V2_TEAM];
  const BlockJob J = B.cjobs[blockIdx.x / HFV_V2_SUB];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (D.vmesh != 2 || !E.stepping) return;
  const int team = threadIdx.x / HFV_V2_TEAM, lane = threadIdx.x % HFV_V2_TEAM;
  const int tiw = (threadIdx.x & 63) / HFV_V2_TEAM;
  const int64_t gi = (int64_t)J.lo + (blockIdx.x % HFV_V2_SUB) * (HFV_BLOCK / HFV_V2_TEAM) + team;
  if (gi >= J.hi) return;  // whole teams; no block barrier follows
  const int64_t N = B.C.n, cb = D.cell_base;
  const int32_t i = (int32_t)(gi - cb);
  if (!B.C.act[gi]) return;  // unbuilt in a partial build (femit stays 0)
  if (WIDE && V.femit[cb + s + i] != HFV_V2_WIDE) return;
  const double xi = B.C.loc[gi], yi = B.C.loc[N + gi], zi = B.C.loc[2 * N + gi];
  V2TeamPoly<TL>& P = polys[team];
  int m = 0, r = 1;
  double r2 = 0.;
  const int kmax = max(6, max(D.vnb[0], D.vnb[1]));
  for (int K = 2; K <= kmax && r == 1; ++K)
    r = v2_team_build<!WIDE, TL>(B, D, V, s, i, xi, yi, zi, K, P, m, r2, lane, tiw);
  if (lane != 0) return;
  if (r == 2 && !WIDE) {
    V.femit[cb + s + i] = HFV_V2_WIDE;
  } else if (r != 0) {
    V.femit[cb + s + i] = 0;
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
  } else {
    v2_finish(B, V, D, E, s, gi, i, xi, yi, zi, P, m);
  }
}

// ---------------------------------------------------------------- faces and CSR
__global__ void __launch_bounds__(HFV_BLOCK) k_v2_faces(Batch B, V2Bufs V) {
  const BlockJob J = B.fjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  const SimEngine& E = B.eng[s];
  if (!D.vmesh || !E.stepping) return;
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
  F.nrm[2 * NF + f] = D.vmesh == 3 ? r.nz : 0.;
  F.fcen[f] = B.C.loc[gi] + r.cx;
  F.fcen[NF + f] = B.C.loc[N + gi] + r.cy;
  F.fcen[2 * NF + f] = D.vmesh == 3 ? B.C.loc[2 * N + gi] + r.cz : B.C.loc[2 * N + gi];
  F.left[f] = (int32_t)gi;
  F.right[f] = r.nb < 0 ? -1 : (int32_t)(cb + r.nb);
  F.shc[f] = (uint8_t)r.code;
  atomicAdd(&V.ccnt[cb + s + lo], 1);
  if (r.nb >= 0) atomicAdd(&V.ccnt[cb + s + r.nb], 1);
}

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_csr_fill(Batch B, V2Bufs V) {
  const BlockJob J = B.fjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  const SimEngine& E = B.eng[s];
  if (!D.vmesh || !E.stepping || E.err) return;
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

#define HFV_CSR_SORT_LDS 24  // a cell's CSR entries sorted in LDS (more: in place in memory)

__global__ void __launch_bounds__(HFV_BLOCK) k_v2_csr_sort(Batch B, V2Bufs V) {
  __shared__ uint64_t ent[HFV_CSR_SORT_LDS][HFV_BLOCK];  // [slot][thread]: (CSR value << 32) | neighbour
  const BlockJob J = B.cjobs[blockIdx.x];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (!D.vmesh || !E.stepping || E.err) return;
  const int64_t gi = (int64_t)J.lo + threadIdx.x;
  if (gi >= J.hi) return;  // (no block barrier follows)
  const int32_t e0 = B.F.off[gi], e1 = B.F.off[gi + 1];
  int32_t* c = B.F.conn;
  int32_t* nb = B.F.nbr;
  const int m = e1 - e0;
  if (m <= HFV_CSR_SORT_LDS) {
    // insertion sort of the cell's entries by face index (then side): CSR values are distinct and
    // non-negative, so the packed words order as they do
    uint64_t* v = &ent[0][threadIdx.x];
    for (int k = 0; k < m; ++k) v[k * HFV_BLOCK] = ((uint64_t)(uint32_t)c[e0 + k] << 32) | (uint32_t)nb[e0 + k];
    for (int k = 1; k < m; ++k) {
      const uint64_t x = v[k * HFV_BLOCK];
      int j = k - 1;
      while (j >= 0 && v[j * HFV_BLOCK] > x) {
        v[(j + 1) * HFV_BLOCK] = v[j * HFV_BLOCK];
        --j;
      }
      v[(j + 1) * HFV_BLOCK] = x;
    }
    for (int k = 0; k < m; ++k) {
      const uint64_t x = v[k * HFV_BLOCK];
      c[e0 + k] = (int32_t)(x >> 32);
      nb[e0 + k] = (int32_t)(uint32_t)x;
    }
    return;
  }
  for (int32_t k = e0 + 1; k < e1; ++k) {  // insertion sort in place
    const int32_t x = c[k], y = nb[k];
    int32_t j = k - 1;
    while (j >= e0 && c[j] > x) {
      c[j + 1] = c[j];
      nb[j + 1] = nb[j];
      --j;
    }
    c[j + 1] = x;
    nb[j + 1] = y;
  }
}

}  // namespace hfv
