// 3D moving meshes rebuilt on the device: Space::volume_calculation (src/space.rs:236-263) for 3D, the
// partial build of the step's active cells (Voronoi::build_partial).  Included by hfv_api.hip after
// hfv_voronoi2d.hpp, whose bin grid, prefix sums, face and CSR assembly it shares.
//
// The cell is the host tessellator's (hydro_rs_amd/csrc/host/voronoi.cpp, Cell3) evaluated with the same
// operations in the same order: a convex polyhedron stored by its dual, every vertex a triangle of three
// plane ids (the initial box's 8 corners oriented by det(n_a, n_b, n_c) > 0), clipped by the bisectors of
// the generators of the (2K+1)^3 surrounding bins in (squared distance, index, image code) order until the
// security radius holds.  A clip removes the vertices beyond the plane (by more than HFV_CLIP_TOL of its
// offset), adds one vertex per boundary edge of the removed region (in the order of the removed vertices
// and their edges; where the three planes meet in no point, the edge's crossing of the plane), keeps the
// others in order and appends the new ones.  The faces are the planes' vertex fans chained into polygons;
// area vector, centroid and the cell's volume and centroid are the host's triangle-fan sums in its order.
//
// One wavefront (a block of its own) builds one cell, everything in LDS: the block's candidates are
// gathered once per K and ranked by their key (each lane ranks every 64th candidate against all), then
// taken in rank order; a clip reads each vertex once (two per lane), lists the removed ones with two
// ballots, compacts the kept ones from registers and makes the new ones one boundary edge per lane; the
// largest vertex radius is a DPP reduction over the lanes' kept and new vertices.  The finish builds each
// plane's fan and polygon on its own lane, with the face sums and the cell's tetrahedron terms, which one
// lane then adds in the host's order.  Blocks of more than HFV_V3_CAP candidates are taken in bands of
// squared distance that fit (chosen from an LDS histogram); only a band that cannot be narrowed enough
// takes each next candidate as a wave minimum over a rescan of the bins.
#pragma once

namespace hfv {

#define HFV_V3_MAXP 64   // planes of a cell: 6 walls and the bisectors that cut it
#define HFV_V3_MAXT 128  // vertices (dual triangles)
#define HFV_V3_FAN 24    // vertices of one face
#define HFV_V3_CAP 192   // candidates held in LDS (a 5^3-bin block at about one generator per bin)
#define HFV_V3_OWN ((HFV_V3_CAP + 63) / 64)
#define HFV_V3_NTRI 192  // fan triangles of a cell (2 V - 4: up to 98 vertices)
static_assert(HFV_V3_FAN % 8 == 0, "fans are searched 8 entries a load");

struct __align__(16) V3Key {
  double d2;
  int32_t j, k;
};

// while clipping: the clip scratch (the removed vertices' indices and plane ids) and the block's
// candidates
struct V3Build {
  uint32_t remp[HFV_V3_MAXT];
  uint8_t remi[HFV_V3_MAXT];
  V3Key key[HFV_V3_CAP];
  double rx[HFV_V3_CAP], ry[HFV_V3_CAP], rz[HFV_V3_CAP];
  uint16_t ord[HFV_V3_CAP];  // candidate of each rank
  uint32_t hist[64];         // squared-distance histogram of a block too large for one pass
  uint32_t nband;            // candidates of the current band
};

// at the finish: per plane its polygon; its fan (u, v, vertex) in vertex order until the polygons are
// chained, then in the same place the cell's tetrahedron terms (volume, centroid numerators) in plane and
// fan order
struct V3Fans {
  uint8_t fu[HFV_V3_MAXP][HFV_V3_FAN], fv[HFV_V3_MAXP][HFV_V3_FAN], ft[HFV_V3_MAXP][HFV_V3_FAN];
};
struct V3Terms {
  double qv[HFV_V3_NTRI], qx[HFV_V3_NTRI], qy[HFV_V3_NTRI], qz[HFV_V3_NTRI];
};
struct V3Fin {
  uint8_t poly[HFV_V3_MAXP][HFV_V3_FAN];
  union {
    V3Fans a;
    V3Terms b;
  } u;
};

struct V3Cell {
  double px[HFV_V3_MAXP], py[HFV_V3_MAXP], pz[HFV_V3_MAXP], pd[HFV_V3_MAXP];
  int32_t pnb[HFV_V3_MAXP], pcode[HFV_V3_MAXP];
  double tx[HFV_V3_MAXT], ty[HFV_V3_MAXT], tz[HFV_V3_MAXT];
  uint32_t tp[HFV_V3_MAXT];  // plane ids a | b << 8 | c << 16
  union {
    V3Build b;
    V3Fin f;
  } u;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the host's V3 helpers: dot = (x x' + y y') + z z', cross componentwise
__device__ __forceinline__ double v3dot(double ax, double ay, double az, double bx, double by, double bz) {
  return ax * bx + ay * by + az * bz;
}

// intersect3 of planes a, b and (cn, cd): false where they meet in no point
__device__ __forceinline__ bool v3_intersect(const V3Cell& C, int a, int b, double cnx, double cny, double cnz,
                                             double cd, double& ox, double& oy, double& oz) {
  const double anx = C.px[a], any = C.py[a], anz = C.pz[a], ad = C.pd[a];
  const double bnx = C.px[b], bny = C.py[b], bnz = C.pz[b], bd = C.pd[b];
  const double bcx = bny * cnz - bnz * cny, bcy = bnz * cnx - bnx * cnz, bcz = bnx * cny - bny * cnx;
  const double det = v3dot(anx, any, anz, bcx, bcy, bcz);
  if (det == 0.0 || !isfinite(det)) return false;
  const double cax = cny * anz - cnz * any, cay = cnz * anx - cnx * anz, caz = cnx * any - cny * anx;
  const double abx = any * bnz - anz * bny, aby = anz * bnx - anx * bnz, abz = anx * bny - any * bnx;
  const double rx = (ad * bcx + bd * cax) + cd * abx, ry = (ad * bcy + bd * cay) + cd * aby,
               rz = (ad * bcz + bd * caz) + cd * abz;
  const double inv = 1.0 / det;
  ox = inv * rx;
  oy = inv * ry;
  oz = inv * rz;
  return true;
}

template <int CTRL, int ROW, int BANK>
__device__ __forceinline__ double dpp_max_step(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const double t = __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW, BANK, false),
                                    __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW, BANK, false));
  return v < t ? t : v;
}

// the maximum over the wavefront (no NaN): within rows by row_shr 1, 2, 4, 8, across rows by
// row_bcast 15 and 31, which leave it in lane 63
__device__ __forceinline__ double wave_max(double v) {
  v = dpp_max_step<0x111, 0xf, 0xf>(v);
  v = dpp_max_step<0x112, 0xf, 0xf>(v);
  v = dpp_max_step<0x114, 0xf, 0xf>(v);
  v = dpp_max_step<0x118, 0xf, 0xf>(v);
  v = dpp_max_step<0x142, 0xa, 0xf>(v);
  v = dpp_max_step<0x143, 0xc, 0xf>(v);
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Cell3::rmax2
__device__ __forceinline__ double v3_rmax2(const V3Cell& C, int nt) {
  const int lane = __lane_id();
  double r = 0.;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = lane + 64 * h;
    if (t < nt) {
      const double q = v3dot(C.tx[t], C.ty[t], C.tz[t], C.tx[t], C.ty[t], C.tz[t]);
      r = r < q ? q : r;
    }
  }
  return wave_max(r);
}

// Cell3::clip by plane (n, d, nb, code): 0 no cut, 1 cut (r2 the new Cell3::rmax2), -1 more than the LDS
// bounds hold.  Every vertex is read once, at the start: the kept ones are compacted from registers, the
// removed ones' plane ids and positions go to the scratch, and the new ones are written in place.
__device__ int v3_clip(V3Cell& C, int& np, int& nt, double& r2, double nx, double ny, double nz, double d,
                       int32_t nb, int32_t code) {
  const int lane = __lane_id();
  V3Build& S = C.u.b;
  // this lane's vertices (lane, lane + 64) and whether the plane removes them
  const bool h0 = lane < nt, h1 = lane + 64 < nt;
  double x0 = 0., y0 = 0., z0 = 0., x1 = 0., y1 = 0., z1 = 0.;
  uint32_t p0 = 0, p1 = 0;
  if (h0) {
    x0 = C.tx[lane];
    y0 = C.ty[lane];
    z0 = C.tz[lane];
    p0 = C.tp[lane];
  }
  if (h1) {
    x1 = C.tx[lane + 64];
    y1 = C.ty[lane + 64];
    z1 = C.tz[lane + 64];
    p1 = C.tp[lane + 64];
  }
  const double dth = d + HFV_CLIP_TOL * d;  // (beyond by more than a rounding error)
  const bool o0 = h0 && v3dot(nx, ny, nz, x0, y0, z0) > dth;
  const bool o1 = h1 && v3dot(nx, ny, nz, x1, y1, z1) > dth;
  const uint64_t m0 = __ballot(o0), m1 = __ballot(o1);
  const int nrem = __popcll(m0) + __popcll(m1);
  if (nrem == 0 || nrem == nt) return 0;
  if (np + 1 > HFV_V3_MAXP) return -1;
  const uint64_t below = (1ull << lane) - 1ull;
  const int r0 = __popcll(m0 & below), r1 = __popcll(m0) + __popcll(m1 & below);
  if (o0) {
    S.remp[r0] = p0;
    S.remi[r0] = (uint8_t)lane;
  }
  if (o1) {
    S.remp[r1] = p1;
    S.remi[r1] = (uint8_t)(lane + 64);
  }
  wave_sync();
  // the kept vertices, in order
  const int nkeep = nt - nrem;
  double q = 0.;  // this lane's largest squared vertex radius
  if (h0 && !o0) {
    const int t = lane - r0;
    C.tx[t] = x0;
    C.ty[t] = y0;
    C.tz[t] = z0;
    C.tp[t] = p0;
    q = v3dot(x0, y0, z0, x0, y0, z0);
  }
  if (h1 && !o1) {
    const int t = lane + 64 - r1;
    C.tx[t] = x1;
    C.ty[t] = y1;
    C.tz[t] = z1;
    C.tp[t] = p1;
    const double q1 = v3dot(x1, y1, z1, x1, y1, z1);
    q = q < q1 ? q1 : q;
  }
  const int pid = np;
  // boundary edges of the removed region, in (removed vertex, edge) order: one new vertex each, after
  // the kept ones
  int nnew = 0;
  for (int q0 = 0; q0 < 3 * nrem; q0 += 64) {
    const int e3 = q0 + lane;
    bool bnd = false, fb = false;
    int x = 0, y = 0, rq = 0;
    double ox = 0., oy = 0., oz = 0.;
    if (e3 < 3 * nrem) {
      rq = e3 / 3;
      const int e = e3 - 3 * rq;
      const uint32_t w = S.remp[rq];
      const int a0 = w & 0xFF, a1 = (w >> 8) & 0xFF, a2 = w >> 16;
      x = e == 0 ? a0 : e == 1 ? a1 : a2;
      y = e == 0 ? a1 : e == 1 ? a2 : a0;
      // internal: another removed vertex has the edge (y, x) (loads batched, no early exit)
      bool internal = false;
      for (int k = 0; k < nrem; k += 4) {
        uint32_t u[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) u[v] = k + v < nrem ? S.remp[k + v] : 0xFFFFFFFFu;  // matches no plane id
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int u0 = u[v] & 0xFF, u1 = (u[v] >> 8) & 0xFF, u2 = u[v] >> 16;
          internal = internal | ((u0 == y) & (u1 == x)) | ((u1 == y) & (u2 == x)) | ((u2 == y) & (u0 == x));
        }
      }
      if (!internal) {
        bnd = true;
        fb = !v3_intersect(C, x, y, nx, ny, nz, d, ox, oy, oz);
      }
    }
    // (rare) the three planes meet in no point: where the edge (x, y) from the removed vertex to the vertex
    // across it (the first with the edge (y, x), a kept one) crosses the plane; the removed vertex's
    // position where that is no crossing.  Each such lane in turn, the others searching their vertices.
    for (uint64_t todo = __ballot(fb); todo; todo &= todo - 1) {
      const int L = __builtin_ctzll(todo);
      const int xq = __builtin_amdgcn_readlane(x, L), yq = __builtin_amdgcn_readlane(y, L);
      const int ri = __builtin_amdgcn_readlane(fb ? (int)S.remi[rq] : 0, L);
      auto has = [&](uint32_t w) {
        const int u0 = w & 0xFF, u1 = (w >> 8) & 0xFF, u2 = w >> 16;
        return ((u0 == yq) & (u1 == xq)) | ((u1 == yq) & (u2 == xq)) | ((u2 == yq) & (u0 == xq));
      };
      const uint64_t k0 = __ballot(h0 && has(p0)), k1 = __ballot(h1 && has(p1));
      const int rs = ri & 63;  // the removed vertex, from the lane that read it
      const double tx = ri < 64 ? __shfl(x0, rs, 64) : __shfl(x1, rs, 64);
      const double ty = ri < 64 ? __shfl(y0, rs, 64) : __shfl(y1, rs, 64);
      const double tz = ri < 64 ? __shfl(z0, rs, 64) : __shfl(z1, rs, 64);
      double cx = tx, cy = ty, cz = tz;
      if (k0 | k1) {
        const int us = k0 ? __builtin_ctzll(k0) : __builtin_ctzll(k1);
        const double ux = k0 ? __shfl(x0, us, 64) : __shfl(x1, us, 64);
        const double uy = k0 ? __shfl(y0, us, 64) : __shfl(y1, us, 64);
        const double uz = k0 ? __shfl(z0, us, 64) : __shfl(z1, us, 64);
        const double a = v3dot(nx, ny, nz, ux, uy, uz) - d, b = v3dot(nx, ny, nz, tx, ty, tz) - d;
        if (a <= 0.0 && b > a) {
          const double sc = a / (a - b);
          cx = ux + sc * (tx - ux);
          cy = uy + sc * (ty - uy);
          cz = uz + sc * (tz - uz);
        }
      }
      if (lane == L) {
        ox = cx;
        oy = cy;
        oz = cz;
      }
    }
    const uint64_t mb = __ballot(bnd);
    const int t = nkeep + nnew + __popcll(mb & below);
    if (bnd && t < HFV_V3_MAXT) {
      C.tx[t] = ox;
      C.ty[t] = oy;
      C.tz[t] = oz;
      C.tp[t] = (uint32_t)x | ((uint32_t)y << 8) | ((uint32_t)pid << 16);
      const double qn = v3dot(ox, oy, oz, ox, oy, oz);
      q = q < qn ? qn : q;
    }
    nnew += __popcll(mb);
  }
  if (nkeep + nnew > HFV_V3_MAXT) return -1;
  if (lane == 0) {
    C.px[pid] = nx;
    C.py[pid] = ny;
    C.pz[pid] = nz;
    C.pd[pid] = d;
    C.pnb[pid] = nb;
    C.pcode[pid] = code;
  }
  r2 = wave_max(q);
  wave_sync();
  np = pid + 1;
  nt = nkeep + nnew;
  return 1;
}

struct V3Cand {
  double d2, rx, ry, rz;
  int32_t j, k, code;
};

__device__ __forceinline__ bool v3_less(const V3Cand& a, const V3Cand& b) { return v2_less(a.d2, a.j, a.k, b.d2, b.j, b.k); }

__device__ __forceinline__ int32_t v3_code_of(int32_t kk) {
  const int s0 = (int8_t)((kk >> 16) & 0xFF), s1 = (int8_t)((kk >> 8) & 0xFF), s2 = (int8_t)(kk & 0xFF);
  return (s0 == 1 ? 1 : s0 == -1 ? 2 : 0) | ((s1 == 1 ? 1 : s1 == -1 ? 2 : 0) << 2) |
         ((s2 == 1 ? 1 : s2 == -1 ? 2 : 0) << 4);
}

// this lane's bins of the (2K+1)^3 block (every 64th): f(bin key, shift codes, whether it is the
// generator's own bin unshifted)
template <class Fn>
__device__ __forceinline__ void v3_lane_bins(const SimDesc& D, double xi, double yi, double zi, int K, Fn&& f) {
  const int lane = __lane_id();
  const bool per = D.periodic != 0;
  const int bi[3] = {v2_bin(xi, D.vw[0], D.vnb[0]), v2_bin(yi, D.vw[1], D.vnb[1]), v2_bin(zi, D.vw[2], D.vnb[2])};
  const int side = 2 * K + 1;
  for (int bb = lane; bb < side * side * side; bb += 64) {
    const int dd[3] = {bb / (side * side) - K, (bb / side) % side - K, bb % side - K};
    int b[3], sc[3];
    bool skip = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      b[a] = bi[a] + dd[a];
      sc[a] = 0;
      if (b[a] < 0 || b[a] >= D.vnb[a]) {
        if (!per) {
          skip = true;
        } else {
          if (b[a] < 0) {
            b[a] += D.vnb[a];
            sc[a] = -1;
          } else {
            b[a] -= D.vnb[a];
            sc[a] = 1;
          }
          if (b[a] < 0 || b[a] >= D.vnb[a]) skip = true;
        }
      }
    }
    if (skip) continue;
    f((b[0] * D.vnb[1] + b[1]) * D.vnb[2] + b[2], sc[0], sc[1], sc[2], dd[0] == 0 && dd[1] == 0 && dd[2] == 0);
  }
}

// this lane's candidates of the (2K+1)^3 block: f(cand) for each
template <class Fn>
__device__ __forceinline__ void v3_lane_candidates(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i,
                                                   double xi, double yi, double zi, int K, Fn&& f) {
  const int64_t N = B.C.n, cb = D.cell_base;
  const double* L = B.C.loc;
  const int32_t* gst = V.gcnt + cb + s;
  const int32_t* gord = V.gord + cb;
  v3_lane_bins(D, xi, yi, zi, K, [&](int key, int s0, int s1, int s2, bool) {
    const int32_t kk = ((s0 & 0xFF) << 16) | ((s1 & 0xFF) << 8) | (s2 & 0xFF);  // the host's memcmp order
    const int32_t code = v3_code_of(kk);
    for (int32_t q = gst[key]; q < gst[key + 1]; ++q) {
      const int32_t j = gord[q];
      if (j == i && s0 == 0 && s1 == 0 && s2 == 0) continue;
      const int64_t gj = cb + j;
      V3Cand c;
      c.rx = (L[gj] + (double)s0 * D.cfg.box[0]) - xi;
      c.ry = (L[N + gj] + (double)s1 * D.cfg.box[1]) - yi;
      c.rz = (L[2 * N + gj] + (double)s2 * D.cfg.box[2]) - zi;
      c.d2 = v3dot(c.rx, c.ry, c.rz, c.rx, c.ry, c.rz);
      c.j = j;
      c.k = kk;
      c.code = code;
      f(c);
    }
  });
}

// Ranks the n candidates in LDS by (squared distance, index, image code): ord[rank] = candidate.
__device__ void v3_rank(V3Cell& C, int n) {
  const int lane = __lane_id();
  V3Build& Q = C.u.b;
  n = __builtin_amdgcn_readfirstlane(n);
  // rank of each of this lane's candidates: the number of keys below it (keys are distinct).  A
  // non-negative double orders as its bits, so keys compare as integers.
  const int nown = __builtin_amdgcn_readfirstlane((n + 63) >> 6);
  uint64_t od[HFV_V3_OWN];
  int32_t oj[HFV_V3_OWN], ok[HFV_V3_OWN], rk[HFV_V3_OWN];
#pragma unroll
  for (int h = 0; h < HFV_V3_OWN; ++h) {
    const int a = lane + 64 * h;
    const V3Key key = Q.key[a < n ? a : 0];
    od[h] = (uint64_t)__double_as_longlong(key.d2);
    oj[h] = key.j;
    ok[h] = key.k;
    rk[h] = 0;
  }
  // By the squared distance alone first (keys below and keys at most equal; broadcast reads, 8 keys in
  // flight); the whole key only where two squared distances are equal.
  int le[HFV_V3_OWN];
#pragma unroll
  for (int h = 0; h < HFV_V3_OWN; ++h) le[h] = 0;
  int b = 0;
  for (; b + 8 <= n; b += 8) {
    uint64_t db[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) db[u] = (uint64_t)__double_as_longlong(Q.key[b + u].d2);
#pragma unroll
    for (int h = 0; h < HFV_V3_OWN; ++h) {
      if (h >= nown) break;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        rk[h] += db[u] < od[h] ? 1 : 0;
        le[h] += db[u] <= od[h] ? 1 : 0;
      }
    }
  }
  for (; b < n; ++b) {
    const uint64_t db = (uint64_t)__double_as_longlong(Q.key[b].d2);
#pragma unroll
    for (int h = 0; h < HFV_V3_OWN; ++h) {
      if (h >= nown) break;
      rk[h] += db < od[h] ? 1 : 0;
      le[h] += db <= od[h] ? 1 : 0;
    }
  }
  bool tie = false;
#pragma unroll
  for (int h = 0; h < HFV_V3_OWN; ++h) tie = tie | ((h < nown) & (lane + 64 * h < n) & (le[h] - rk[h] > 1));
  if (__ballot(tie)) {
#pragma unroll
    for (int h = 0; h < HFV_V3_OWN; ++h) rk[h] = 0;
    for (int c = 0; c < n; ++c) {
      const V3Key kb = Q.key[c];
      const uint64_t db = (uint64_t)__double_as_longlong(kb.d2);
#pragma unroll
      for (int h = 0; h < HFV_V3_OWN; ++h) {
        if (h >= nown) break;
        // (bitwise, not short-circuit: no branches)
        const bool lt = (db < od[h]) | ((db == od[h]) & ((kb.j < oj[h]) | ((kb.j == oj[h]) & (kb.k < ok[h]))));
        rk[h] += lt ? 1 : 0;
      }
    }
  }
#pragma unroll
  for (int h = 0; h < HFV_V3_OWN; ++h) {
    const int a = lane + 64 * h;
    if (h < nown && a < n) Q.ord[rk[h]] = (uint16_t)a;
  }
  wave_sync();
}

// The block's candidates into LDS, ranked by (squared distance, index, image code): their count, or -1
// beyond HFV_V3_CAP.
__device__ int v3_gather(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i, double xi, double yi,
                         double zi, int K, V3Cell& C) {
  const int lane = __lane_id();
  const int32_t* gst = V.gcnt + D.cell_base + s;
  int cnt = 0;
  v3_lane_bins(D, xi, yi, zi, K, [&](int key, int, int, int, bool own) { cnt += gst[key + 1] - gst[key] - (own ? 1 : 0); });
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const int n = __builtin_amdgcn_readlane(incl, 63);
  if (n > HFV_V3_CAP) return -1;
  V3Build& Q = C.u.b;
  int at = incl - cnt;
  v3_lane_candidates(B, D, V, s, i, xi, yi, zi, K, [&](const V3Cand& c) {
    if (at >= n) return;  // (the counts agree: the generator's own entry is the one skipped)
    V3Key key;
    key.d2 = c.d2;
    key.j = c.j;
    key.k = c.k;
    Q.key[at] = key;
    Q.rx[at] = c.rx;
    Q.ry[at] = c.ry;
    Q.rz[at] = c.rz;
    ++at;
  });
  wave_sync();
  v3_rank(C, n);
  return n;
}

// The bin of squared distance d2 among 64 of width step from lo (the last one open), by exact comparisons
// with the edges lo + k step, so that a band [lo, lo + (b + 1) step) holds exactly bins 0..b.
__device__ __forceinline__ int v3_band_bin(double d2, double lo, double step) {
  int k = (int)fmin(fmax((d2 - lo) / step, 0.), 63.);
  while (k < 63 && d2 >= lo + (double)(k + 1) * step) ++k;
  while (k > 0 && d2 < lo + (double)k * step) --k;
  return k;
}

// A block of more than HFV_V3_CAP candidates, taken in bands of squared distance: the candidates of the
// next band [lo, hi) (hi chosen from a histogram so that at most HFV_V3_CAP fall in it; +inf for the
// last) gathered into LDS and ranked.  Returns their count, or -1 where even a narrow band near lo holds
// more than HFV_V3_CAP.
__device__ int v3_band(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i, double xi, double yi,
                       double zi, int K, V3Cell& C, double lo, double& hi) {
  const int lane = __lane_id();
  V3Build& Q = C.u.b;
  double dmax = 0.;  // beyond every candidate of the block
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double r = (K + 1) * D.vw[a];
    dmax += r * r;
  }
  if (!(dmax > lo)) dmax = 2. * lo + 1.;
  int b = -1;
  for (int zoom = 0; zoom < 4 && b < 0; ++zoom) {
    const double step = (dmax - lo) / 64.;
    Q.hist[lane] = 0u;
    wave_sync();
    v3_lane_candidates(B, D, V, s, i, xi, yi, zi, K, [&](const V3Cand& c) {
      if (c.d2 >= lo) atomicAdd(&Q.hist[v3_band_bin(c.d2, lo, step)], 1u);
    });
    wave_sync();
    int incl = (int)Q.hist[lane];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    b = __popcll(__ballot(incl <= HFV_V3_CAP)) - 1;  // the longest prefix of bins that fits
    if (b >= 0) hi = b == 63 ? (double)INFINITY : lo + (double)(b + 1) * step;
    if (b >= 0 && !(hi > lo)) b = -1;  // (a band too narrow for the doubles)
    dmax = lo + step;  // else: the first bin alone holds too many; split it
  }
  if (b < 0) return -1;
  if (lane == 0) Q.nband = 0u;
  wave_sync();
  v3_lane_candidates(B, D, V, s, i, xi, yi, zi, K, [&](const V3Cand& c) {
    if (!(c.d2 >= lo && c.d2 < hi)) return;
    const uint32_t at = atomicAdd(&Q.nband, 1u);  // (any order: the rank orders them)
    if (at >= HFV_V3_CAP) return;                 // (cannot happen: the histogram counted them)
    V3Key key;
    key.d2 = c.d2;
    key.j = c.j;
    key.k = c.k;
    Q.key[at] = key;
    Q.rx[at] = c.rx;
    Q.ry[at] = c.ry;
    Q.rz[at] = c.rz;
  });
  wave_sync();
  const int n = __builtin_amdgcn_readfirstlane((int)min(Q.nband, (uint32_t)HFV_V3_CAP));
  v3_rank(C, n);
  return n;
}

__device__ __forceinline__ void wave_min_cand(bool& have, V3Cand& b) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int oh = __shfl_xor((int)have, o, 64);
    V3Cand c;
    c.d2 = __shfl_xor(b.d2, o, 64);
    c.rx = __shfl_xor(b.rx, o, 64);
    c.ry = __shfl_xor(b.ry, o, 64);
    c.rz = __shfl_xor(b.rz, o, 64);
    c.j = __shfl_xor(b.j, o, 64);
    c.k = __shfl_xor(b.k, o, 64);
    c.code = __shfl_xor(b.code, o, 64);
    if (oh && (!have || v3_less(c, b))) {
      have = true;
      b = c;
    }
  }
}

// The cell for candidate bins within K: 0 clipped (finish next), 1 retry with a larger K, 3 beyond the
// LDS bounds.
__device__ int v3_build(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i, double xi, double yi,
                        double zi, int K, V3Cell& C, int& np, int& nt, double& r2) {
  const int lane = __lane_id();
  const double w0 = D.vw[0], w1 = D.vw[1], w2 = D.vw[2];
  const bool per = D.periodic != 0;
  double reach = 0.;
  reach = reach < K * w0 ? K * w0 : reach;
  reach = reach < K * w1 ? K * w1 : reach;
  reach = reach < K * w2 ? K * w2 : reach;
  const double lo[3] = {per ? -reach : -xi, per ? -reach : -yi, per ? -reach : -zi};
  const double hi[3] = {per ? reach : D.cfg.box[0] - xi, per ? reach : D.cfg.box[1] - yi,
                        per ? reach : D.cfg.box[2] - zi};
  if (lane < 6) {  // walls 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z
    const int a = lane >> 1;
    const bool up = lane & 1;
    C.px[lane] = a == 0 ? (up ? 1. : -1.) : 0.;
    C.py[lane] = a == 1 ? (up ? 1. : -1.) : 0.;
    C.pz[lane] = a == 2 ? (up ? 1. : -1.) : 0.;
    C.pd[lane] = up ? hi[a] : -lo[a];
    C.pnb[lane] = -1 - lane;
    C.pcode[lane] = 0;
  }
  wave_sync();
  if (lane < 8) {  // the cube's corners (sx, sy, sz), oriented so that det(n_a, n_b, n_c) > 0
    const int a = lane >> 2, b = 2 + ((lane >> 1) & 1), c = 4 + (lane & 1);
    const double cx = C.py[b] * C.pz[c] - C.pz[b] * C.py[c], cy = C.pz[b] * C.px[c] - C.px[b] * C.pz[c],
                 cz = C.px[b] * C.py[c] - C.py[b] * C.px[c];
    const double det = v3dot(C.px[a], C.py[a], C.pz[a], cx, cy, cz);
    const int t0 = det > 0 ? a : b, t1 = det > 0 ? b : a;
    C.tp[lane] = (uint32_t)t0 | ((uint32_t)t1 << 8) | ((uint32_t)c << 16);
    double ox = 0., oy = 0., oz = 0.;
    v3_intersect(C, t0, t1, C.px[c], C.py[c], C.pz[c], C.pd[c], ox, oy, oz);
    C.tx[lane] = ox;
    C.ty[lane] = oy;
    C.tz[lane] = oz;
  }
  wave_sync();
  np = 6;
  nt = 8;
  r2 = v3_rmax2(C, nt);
  // the candidates in rank order until the security radius holds: the whole block at once, or in bands
  // of squared distance (each band's candidates all beyond the previous band's)
  const V3Build& Q = C.u.b;
  int n = v3_gather(B, D, V, s, i, xi, yi, zi, K, C);
  double blo = 0., bhi = (double)INFINITY;
  if (n < 0) n = v3_band(B, D, V, s, i, xi, yi, zi, K, C, blo, bhi);
  while (n >= 0) {
    bool stop = false;
    for (int q = 0; q < n; ++q) {
      const int a = __builtin_amdgcn_readfirstlane(Q.ord[q]);
      const V3Key key = Q.key[a];
      if (0.25 * key.d2 > r2) {  // security radius: the bisector lies beyond every vertex
        stop = true;
        break;
      }
      if (v3_clip(C, np, nt, r2, Q.rx[a], Q.ry[a], Q.rz[a], 0.5 * key.d2, key.j, v3_code_of(key.k)) < 0) return 3;
    }
    if (stop || bhi == (double)INFINITY) break;
    blo = bhi;
    n = v3_band(B, D, V, s, i, xi, yi, zi, K, C, blo, bhi);
  }
  if (n < 0) {  // a band too dense for the LDS: from blo on, the next one is a wave minimum over a rescan
    bool have_last = false;
    V3Cand last{};
    while (true) {
      bool have = false;
      V3Cand best{};
      v3_lane_candidates(B, D, V, s, i, xi, yi, zi, K, [&](const V3Cand& c) {
        if (c.d2 >= blo && (!have_last || v3_less(last, c)) && (!have || v3_less(c, best))) {
          have = true;
          best = c;
        }
      });
      wave_min_cand(have, best);
      if (!have || 0.25 * best.d2 > r2) break;
      if (v3_clip(C, np, nt, r2, best.rx, best.ry, best.rz, 0.5 * best.d2, best.j, best.code) < 0) return 3;
      have_last = true;
      last = best;
    }
  }
  // completeness: every unvisited generator lies farther than K bin widths
  double minw = 1e300;
  minw = K * w0 < minw ? K * w0 : minw;
  minw = K * w1 < minw ? K * w1 : minw;
  minw = K * w2 < minw ? K * w2 : minw;
  if (4.0 * r2 > minw * minw) return 1;
  return 0;
}

// Cell3::finish on the clipped cell.  Returns 0, 1 (a periodic cell touching the initial region: retry
// with a larger K) or 3 (a face of more than HFV_V3_FAN vertices); on 0 the faces and the geometry are
// written.
__device__ int v3_finish(const Batch& B, const V2Bufs& V, const SimDesc& D, SimEngine& E, int s, int64_t gi, int32_t i,
                         double xi, double yi, double zi, V3Cell& C, int np, int nt, double r2) {
  const int lane = __lane_id();
  const int p = lane;
  V3Fin& F = C.u.f;  // over the candidates, read for the last time before the completeness test
  wave_sync();
  // plane p's fan: its vertices rotated to (p, u, v), in vertex order; chained into a polygon
  int m = 0;
  bool over = false;
  if (p < np) {
    int f = 0;
    for (int t0 = 0; t0 < nt; t0 += 8) {  // (8 loads in flight)
      uint32_t w[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) w[e] = t0 + e < nt ? C.tp[t0 + e] : 0xFFFFFFFFu;  // matches no plane id
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int a = w[e] & 0xFF, b = (w[e] >> 8) & 0xFF, c = w[e] >> 16;
        const bool ia = a == p, ib = b == p, ic = c == p;
        const int u = ia ? b : ib ? c : ic ? a : -1, v = ia ? c : ib ? a : b;
        if (u >= 0) {
          if (f < HFV_V3_FAN) {
            F.u.a.fu[p][f] = (uint8_t)u;
            F.u.a.fv[p][f] = (uint8_t)v;
            F.u.a.ft[p][f] = (uint8_t)(t0 + e);
          }
          ++f;
        }
      }
    }
    over = f > HFV_V3_FAN;
    if (f >= 3 && !over) {
      // from (u, v) the next entry around p is the first one whose u is this v (8 entries a load)
      const uint2* fu8 = reinterpret_cast<const uint2*>(&F.u.a.fu[p][0]);
      int cur = 0;
      for (int step = 0; step < f; ++step) {
        F.poly[p][m++] = F.u.a.ft[p][cur];
        const uint32_t want = F.u.a.fv[p][cur];
        int nxt = -1;
        for (int q0 = 0; q0 < f; q0 += 8) {
          const uint2 w = fu8[q0 >> 3];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const uint32_t u = ((e < 4 ? w.x : w.y) >> (8 * (e & 3))) & 0xFFu;
            nxt = ((nxt < 0) & (q0 + e < f) & (u == want)) ? q0 + e : nxt;
          }
        }
        if (nxt < 0) break;
        cur = nxt;
        if (cur == 0) break;
      }
    }
    if (m < 3) m = 0;
  }
  if (__ballot(over)) return 3;
  // the plane's fan triangles: their place among the cell's (planes in order), by a wave prefix sum
  const int nq = m >= 3 ? m - 2 : 0;
  int incl = nq;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const int ntri = __builtin_amdgcn_readlane(incl, 63);
  if (ntri > HFV_V3_NTRI) return 3;  // (2 V - 4 on a closed cell)
  const int q0 = incl - nq;
  wave_sync();  // the fans are read no more: their place takes the terms
  // the face's area vector and centroid; the cell's tetrahedron terms (vertex fan with the generator)
  double area = 0., fcx = 0., fcy = 0., fcz = 0.;
  if (m >= 3) {
    const int t0 = F.poly[p][0];
    const double p0x = C.tx[t0], p0y = C.ty[t0], p0z = C.tz[t0];
    double avx = 0., avy = 0., avz = 0., acx = 0., acy = 0., acz = 0., asum = 0.;
    for (int k = 1; k + 1 < m; ++k) {
      const int ta = F.poly[p][k], tb = F.poly[p][k + 1];
      const double Ax = C.tx[ta], Ay = C.ty[ta], Az = C.tz[ta], Bx = C.tx[tb], By = C.ty[tb], Bz = C.tz[tb];
      const double ax = Ax - p0x, ay = Ay - p0y, az = Az - p0z, bx = Bx - p0x, by = By - p0y, bz = Bz - p0z;
      const double crx = ay * bz - az * by, cry = az * bx - ax * bz, crz = ax * by - ay * bx;
      const double a = 0.5 * sqrt(v3dot(crx, cry, crz, crx, cry, crz));
      avx = avx + 0.5 * crx;
      avy = avy + 0.5 * cry;
      avz = avz + 0.5 * crz;
      asum += a;
      const double s3 = a / 3.0;
      acx = acx + s3 * ((p0x + Ax) + Bx);
      acy = acy + s3 * ((p0y + Ay) + By);
      acz = acz + s3 * ((p0z + Az) + Bz);
      const double ux = Ay * Bz - Az * By, uy = Az * Bx - Ax * Bz, uz = Ax * By - Ay * Bx;
      const double tv = fabs(v3dot(p0x, p0y, p0z, ux, uy, uz)) / 6.0;
      const double s4 = tv / 4.0;
      const int qi = q0 + k - 1;
      F.u.b.qv[qi] = tv;
      F.u.b.qx[qi] = s4 * ((p0x + Ax) + Bx);
      F.u.b.qy[qi] = s4 * ((p0y + Ay) + By);
      F.u.b.qz[qi] = s4 * ((p0z + Az) + Bz);
    }
    area = sqrt(v3dot(avx, avy, avz, avx, avy, avz));
    if (asum > 0) {
      const double inv = 1.0 / asum;
      fcx = inv * acx;
      fcy = inv * acy;
      fcz = inv * acz;
    } else {
      fcx = p0x;
      fcy = p0y;
      fcz = p0z;
    }
  }
  const int32_t pnb = p < np ? C.pnb[p] : 0;
  // a periodic cell must not touch the initial region (a wall face of positive area)
  if (D.periodic && __ballot(p < np && pnb < 0 && area > 0.)) return 1;
  const int64_t cb = D.cell_base, N = B.C.n;
  const bool emit = p < np && area > 0. && v2_emits(B, cb, pnb, i);
  const uint64_t me = __ballot(emit);
  const int cnt = __popcll(me);
  int start = 0;
  if (lane == 0) start = atomicAdd(&E.nf_tmp, cnt);
  start = __shfl(start, 0, 64);
  if ((int64_t)start + cnt > D.face_cap) {
    if (lane == 0) {
      V.femit[cb + s + i] = 0;
      atomicCAS(&B.cnt->err, 0, E_MESH);
      atomicCAS(&E.err, 0, E_MESH);
    }
    return 0;
  }
  if (emit) {
    const double inv = 1.0 / sqrt(v3dot(C.px[p], C.py[p], C.pz[p], C.px[p], C.py[p], C.pz[p]));
    V2Face f;
    f.area = area;
    f.cx = fcx;
    f.cy = fcy;
    f.cz = fcz;
    f.nx = inv * C.px[p];
    f.ny = inv * C.py[p];
    f.nz = inv * C.pz[p];
    f.nb = pnb < 0 ? -1 : pnb;
    f.code = C.pcode[p];
    V.vf[D.face_base + start + __popcll(me & ((1ull << lane) - 1ull))] = f;
  }
  wave_sync();
  if (lane != 0) return 0;
  V.vst[gi] = start;
  V.femit[cb + s + i] = cnt;
  // the cell's volume and centroid: the tetrahedron terms added in the host's order
  double vol = 0., cenx = 0., ceny = 0., cenz = 0.;
  for (int q = 0; q < ntri; ++q) {
    vol += F.u.b.qv[q];
    cenx = cenx + F.u.b.qx[q];
    ceny = ceny + F.u.b.qy[q];
    cenz = cenz + F.u.b.qz[q];
  }
  if (B.C.act[gi]) {  // Particle::update_geometry (part.rs:151-158)
    B.C.vol[gi] = vol;
    const double iv = vol > 0. ? 1.0 / vol : 0.;
    B.C.cen[gi] = xi + (vol > 0. ? iv * cenx : 0.);
    B.C.cen[N + gi] = yi + (vol > 0. ? iv * ceny : 0.);
    B.C.cen[2 * N + gi] = zi + (vol > 0. ? iv * cenz : 0.);
    B.C.srad[gi] = 1.5 * 2.0 * sqrt(r2);
  }
  return 0;
}

// grid: HFV_BLOCK one-wavefront blocks per cell chunk, one cell each
__global__ void __launch_bounds__(64) k_v3_cells(Batch B, V2Bufs V) {
  __shared__ V3Cell C;
  const BlockJob J = B.cjobs[blockIdx.x / HFV_BLOCK];
  if (J.sim < 0) return;
  const int s = J.sim;
  const SimDesc& D = B.sims[s];
  SimEngine& E = B.eng[s];
  if (D.vmesh != 3 || !E.stepping) return;
  const int64_t gi = (int64_t)J.lo + blockIdx.x % HFV_BLOCK;
  if (gi >= J.hi) return;
  if (!B.C.act[gi]) return;  // unbuilt in a partial build (femit stays 0)
  const int64_t N = B.C.n, cb = D.cell_base;
  const int32_t i = (int32_t)(gi - cb);
  const double xi = B.C.loc[gi], yi = B.C.loc[N + gi], zi = B.C.loc[2 * N + gi];
  const int kmax = max(6, max(D.vnb[0], max(D.vnb[1], D.vnb[2])));
  int np = 0, nt = 0, r = 1;
  double r2 = 0.;
  for (int K = 2; K <= kmax && r == 1; ++K) {
    r = v3_build(B, D, V, s, i, xi, yi, zi, K, C, np, nt, r2);
    if (r == 0) r = v3_finish(B, V, D, E, s, gi, i, xi, yi, zi, C, np, nt, r2);
  }
  if (r != 0 && __lane_id() == 0) {
    V.femit[cb + s + i] = 0;
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
  }
}

}  // namespace hfv
