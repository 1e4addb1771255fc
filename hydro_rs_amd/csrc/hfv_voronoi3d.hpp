// 3D moving meshes rebuilt on the device: Space::volume_calculation (src/space.rs:236-263) for 3D, the
// partial build of the step's active cells (Voronoi::build_partial).  Included by hfv_api.hip after
// hfv_voronoi2d.hpp, whose bin grid, prefix sums, face and CSR assembly it shares.
//
// The cell is the host tessellator's (hydro_rs_amd/csrc/host/voronoi.cpp, Cell3) evaluated with the same
// operations in the same order: a convex polyhedron stored by its dual, every vertex a triangle of three
// plane ids (the initial box's 8 corners oriented by det(n_a, n_b, n_c) > 0), clipped by the bisectors of
// the generators of the (2K+1)^3 surrounding bins in (squared distance, index, image code) order until the
// security radius holds.  A clip removes the vertices beyond the plane, adds one vertex per boundary edge
// of the removed region (in the order of the removed vertices and their edges), keeps the others in
// order and appends the new ones.  The faces are the planes' vertex fans chained into polygons; area
// vector, centroid and the cell's volume and centroid are the host's triangle-fan sums in its order.
//
// One wavefront (a block of its own) builds one cell: the polyhedron,
// the removed list, the new vertices and the planes' fans and polygons live in LDS, each lane handling
// every 64th vertex, edge or plane; the next candidate is a wave minimum.  k_v3_cells<false> keeps each
// lane's candidates in registers (at most HFV_V3_TC); cells that need more go to k_v3_cells<true>, which
// rescans the bins for every candidate.
#pragma once

namespace hfv {

#define HFV_V3_MAXP 64   // planes of a cell: 6 walls and the bisectors that cut it
#define HFV_V3_MAXT 128  // vertices (dual triangles)
#define HFV_V3_FAN 24    // vertices of one face
#define HFV_V3_TC 4

struct V3Cell {
  double px[HFV_V3_MAXP], py[HFV_V3_MAXP], pz[HFV_V3_MAXP], pd[HFV_V3_MAXP];
  int32_t pnb[HFV_V3_MAXP], pcode[HFV_V3_MAXP];
  double tx[HFV_V3_MAXT], ty[HFV_V3_MAXT], tz[HFV_V3_MAXT];
  uint8_t ta[HFV_V3_MAXT], tb[HFV_V3_MAXT], tc[HFV_V3_MAXT];
  // clip scratch: the removed vertices, the new vertices
  uint8_t rem[HFV_V3_MAXT];
  double sx[HFV_V3_MAXT], sy[HFV_V3_MAXT], sz[HFV_V3_MAXT];
  uint8_t sa[HFV_V3_MAXT], sb[HFV_V3_MAXT];
  // finish: per plane its fan (u, v, vertex) in vertex order, then its polygon (vertex ids)
  uint8_t fu[HFV_V3_MAXP][HFV_V3_FAN], fv[HFV_V3_MAXP][HFV_V3_FAN], ft[HFV_V3_MAXP][HFV_V3_FAN];
  uint8_t fn[HFV_V3_MAXP], pn[HFV_V3_MAXP];
  uint8_t poly[HFV_V3_MAXP][HFV_V3_FAN];
  double farea[HFV_V3_MAXP], fcx[HFV_V3_MAXP], fcy[HFV_V3_MAXP], fcz[HFV_V3_MAXP];
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

// intersect3 of planes a, b, c: false where they meet in no point
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

__device__ __forceinline__ double wave_max(double r) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double q = __shfl_xor(r, o, 64);
    r = r < q ? q : r;
  }
  return r;
}

__device__ __forceinline__ double v3_rmax2(const V3Cell& C, int nt) {
  const int lane = __lane_id();
  double r = 0.;
  for (int t = lane; t < nt; t += 64) {
    const double q = v3dot(C.tx[t], C.ty[t], C.tz[t], C.tx[t], C.ty[t], C.tz[t]);
    r = r < q ? q : r;
  }
  return wave_max(r);
}

// Cell3::clip by plane (n, d, nb, code): 0 no cut, 1 cut, -1 more than the LDS bounds hold.
__device__ int v3_clip(V3Cell& C, int& np, int& nt, double nx, double ny, double nz, double d, int32_t nb, int32_t code) {
  const int lane = __lane_id();
  // the removed vertices (ascending), as a 128-bit mask
  const bool o0 = lane < nt && v3dot(nx, ny, nz, C.tx[lane], C.ty[lane], C.tz[lane]) > d;
  const bool o1 = lane + 64 < nt && v3dot(nx, ny, nz, C.tx[lane + 64], C.ty[lane + 64], C.tz[lane + 64]) > d;
  const uint64_t m0 = __ballot(o0), m1 = __ballot(o1);
  const int nrem = __popcll(m0) + __popcll(m1);
  if (nrem == 0 || nrem == nt) return 0;
  if (np + 1 > HFV_V3_MAXP) return -1;
  const uint64_t below = (1ull << lane) - 1ull;
  if (o0) C.rem[__popcll(m0 & below)] = (uint8_t)lane;
  if (o1) C.rem[__popcll(m0) + __popcll(m1 & below)] = (uint8_t)(lane + 64);
  wave_sync();
  const int pid = np;
  // boundary edges of the removed region, in (removed vertex, edge) order: one new vertex each
  int nnew = 0;
  for (int q0 = 0; q0 < 3 * nrem; q0 += 64) {
    const int q = q0 + lane;
    bool bnd = false;
    int x = 0, y = 0;
    double ox = 0., oy = 0., oz = 0.;
    if (q < 3 * nrem) {
      const int ri = C.rem[q / 3], e = q % 3;
      const int p0 = C.ta[ri], p1 = C.tb[ri], p2 = C.tc[ri];
      x = e == 0 ? p0 : e == 1 ? p1 : p2;
      y = e == 0 ? p1 : e == 1 ? p2 : p0;
      bool internal = false;
      for (int k = 0; k < nrem && !internal; ++k) {
        const int u = C.rem[k];
        const int u0 = C.ta[u], u1 = C.tb[u], u2 = C.tc[u];
        internal = (u0 == y && u1 == x) || (u1 == y && u2 == x) || (u2 == y && u0 == x);
      }
      if (!internal) {
        bnd = true;
        if (!v3_intersect(C, x, y, nx, ny, nz, d, ox, oy, oz)) {
          ox = C.tx[ri];
          oy = C.ty[ri];
          oz = C.tz[ri];
        }
      }
    }
    const uint64_t mb = __ballot(bnd);
    if (bnd) {
      const int slot = nnew + __popcll(mb & below);
      if (slot < HFV_V3_MAXT) {
        C.sx[slot] = ox;
        C.sy[slot] = oy;
        C.sz[slot] = oz;
        C.sa[slot] = (uint8_t)x;
        C.sb[slot] = (uint8_t)y;
      }
    }
    nnew += __popcll(mb);
  }
  const int nkeep = nt - nrem;
  if (nkeep + nnew > HFV_V3_MAXT) return -1;
  // compact the kept vertices in order (staged through registers), then append the new ones
  double kx[2] = {0., 0.}, ky[2] = {0., 0.}, kz[2] = {0., 0.};
  int ka[2] = {0, 0}, kb[2] = {0, 0}, kc[2] = {0, 0}, kd[2] = {-1, -1};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = lane + 64 * h;
    const bool out = h == 0 ? o0 : o1;
    if (t < nt && !out) {
      const int before = h == 0 ? __popcll(m0 & below) : __popcll(m0) + __popcll(m1 & below);
      kd[h] = t - before;
      kx[h] = C.tx[t];
      ky[h] = C.ty[t];
      kz[h] = C.tz[t];
      ka[h] = C.ta[t];
      kb[h] = C.tb[t];
      kc[h] = C.tc[t];
    }
  }
  wave_sync();
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (kd[h] >= 0) {
      const int t = kd[h];
      C.tx[t] = kx[h];
      C.ty[t] = ky[h];
      C.tz[t] = kz[h];
      C.ta[t] = (uint8_t)ka[h];
      C.tb[t] = (uint8_t)kb[h];
      C.tc[t] = (uint8_t)kc[h];
    }
  for (int k = lane; k < nnew; k += 64) {
    const int t = nkeep + k;
    C.tx[t] = C.sx[k];
    C.ty[t] = C.sy[k];
    C.tz[t] = C.sz[k];
    C.ta[t] = C.sa[k];
    C.tb[t] = C.sb[k];
    C.tc[t] = (uint8_t)pid;
  }
  if (lane == 0) {
    C.px[pid] = nx;
    C.py[pid] = ny;
    C.pz[pid] = nz;
    C.pd[pid] = d;
    C.pnb[pid] = nb;
    C.pcode[pid] = code;
  }
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

// this lane's candidates of the (2K+1)^3 block: f(cand) for each
template <class Fn>
__device__ __forceinline__ void v3_lane_candidates(const Batch& B, const SimDesc& D, const V2Bufs& V, int s, int32_t i,
                                                   double xi, double yi, double zi, int K, Fn&& f) {
  const int lane = __lane_id();
  const int64_t N = B.C.n, cb = D.cell_base;
  const double* L = B.C.loc;
  const bool per = D.periodic != 0;
  const int bi[3] = {v2_bin(xi, D.vw[0], D.vnb[0]), v2_bin(yi, D.vw[1], D.vnb[1]), v2_bin(zi, D.vw[2], D.vnb[2])};
  const int32_t* gst = V.gcnt + cb + s;
  const int32_t* gord = V.gord + cb;
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
    const int key = (b[0] * D.vnb[1] + b[1]) * D.vnb[2] + b[2];
    const int32_t kk = ((sc[0] & 0xFF) << 16) | ((sc[1] & 0xFF) << 8) | (sc[2] & 0xFF);
    const int32_t code = (sc[0] == 1 ? 1 : sc[0] == -1 ? 2 : 0) | ((sc[1] == 1 ? 1 : sc[1] == -1 ? 2 : 0) << 2) |
                         ((sc[2] == 1 ? 1 : sc[2] == -1 ? 2 : 0) << 4);
    for (int32_t q = gst[key]; q < gst[key + 1]; ++q) {
      const int32_t j = gord[q];
      if (j == i && sc[0] == 0 && sc[1] == 0 && sc[2] == 0) continue;
      const int64_t gj = cb + j;
      V3Cand c;
      c.rx = (L[gj] + (double)sc[0] * D.cfg.box[0]) - xi;
      c.ry = (L[N + gj] + (double)sc[1] * D.cfg.box[1]) - yi;
      c.rz = (L[2 * N + gj] + (double)sc[2] * D.cfg.box[2]) - zi;
      c.d2 = v3dot(c.rx, c.ry, c.rz, c.rx, c.ry, c.rz);
      c.j = j;
      c.k = kk;
      c.code = code;
      f(c);
    }
  }
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

// The cell for candidate bins within K: 0 clipped (finish next), 1 retry with a larger K, 2 too many
// candidates for the registers (LIST), 3 beyond the LDS bounds.
template <bool LIST>
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
    C.ta[lane] = (uint8_t)t0;
    C.tb[lane] = (uint8_t)t1;
    C.tc[lane] = (uint8_t)c;
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
  V3Cand cl[LIST ? HFV_V3_TC : 1];
  int cnt = 0;
  if (LIST) {
    v3_lane_candidates(B, D, V, s, i, xi, yi, zi, K, [&](const V3Cand& c) {
#pragma unroll
      for (int q = 0; q < (LIST ? HFV_V3_TC : 1); ++q)
        if (q == cnt) cl[q] = c;
      ++cnt;
    });
    if (__ballot(cnt > HFV_V3_TC)) return 2;
  }
  bool have_last = false;
  V3Cand last{};
  while (true) {
    bool have = false;
    V3Cand best{};
    auto consider = [&](const V3Cand& c) {
      if ((!have_last || v3_less(last, c)) && (!have || v3_less(c, best))) {
        have = true;
        best = c;
      }
    };
    if (LIST) {
#pragma unroll
      for (int q = 0; q < (LIST ? HFV_V3_TC : 1); ++q)
        if (q < cnt) consider(cl[q]);
    } else {
      v3_lane_candidates(B, D, V, s, i, xi, yi, zi, K, consider);
    }
    wave_min_cand(have, best);
    if (!have || 0.25 * best.d2 > r2) break;  // security radius
    const int c = v3_clip(C, np, nt, best.rx, best.ry, best.rz, 0.5 * best.d2, best.j, best.code);
    if (c < 0) return 3;
    if (c > 0) r2 = v3_rmax2(C, nt);
    have_last = true;
    last = best;
  }
  double minw = 1e300;
  minw = K * w0 < minw ? K * w0 : minw;
  minw = K * w1 < minw ? K * w1 : minw;
  minw = K * w2 < minw ? K * w2 : minw;
  if (4.0 * r2 > minw * minw) return 1;
  return 0;
}

// Cell3::finish on the clipped cell: per plane its fan chained into a polygon (lane per plane), the face
// areas and centroids (lane per plane), the cell's volume and centroid summed over the planes and their
// fans in order (lane 0).  Returns 0, 1 (periodic cell touching the initial region: retry with a larger
// K) or 3 (a fan beyond the LDS bound); on 0 the faces and the geometry are written.
__device__ int v3_finish(const Batch& B, const V2Bufs& V, const SimDesc& D, SimEngine& E, int s, int64_t gi, int32_t i,
                         double xi, double yi, double zi, V3Cell& C, int np, int nt, double r2) {
  const int lane = __lane_id();
  bool over = false;
  for (int p = lane; p < np; p += 64) {
    int f = 0;
    for (int t = 0; t < nt; ++t) {
      const int a = C.ta[t], b = C.tb[t], c = C.tc[t];
      int u = -1, v = -1;
      if (a == p) {
        u = b;
        v = c;
      } else if (b == p) {
        u = c;
        v = a;
      } else if (c == p) {
        u = a;
        v = b;
      }
      if (u < 0) continue;
      if (f < HFV_V3_FAN) {
        C.fu[p][f] = (uint8_t)u;
        C.fv[p][f] = (uint8_t)v;
        C.ft[p][f] = (uint8_t)t;
      }
      ++f;
    }
    if (f > HFV_V3_FAN) over = true;
    C.fn[p] = (uint8_t)(f > HFV_V3_FAN ? 0 : f);
    // chain: from (u, v) the next fan entry around p is the first one whose u is this v
    int m = 0;
    if (f >= 3 && f <= HFV_V3_FAN) {
      int cur = 0;
      for (int step = 0; step < f; ++step) {
        C.poly[p][m++] = C.ft[p][cur];
        const int want = C.fv[p][cur];
        int nxt = -1;
        for (int q = 0; q < f; ++q)
          if (C.fu[p][q] == want) {
            nxt = q;
            break;
          }
        if (nxt < 0) break;
        cur = nxt;
        if (cur == 0) break;
      }
    }
    C.pn[p] = (uint8_t)(m >= 3 ? m : 0);
    // the face: area vector and centroid of its triangle fan
    double area = 0., cx = 0., cy = 0., cz = 0.;
    if (m >= 3) {
      const int t0 = C.poly[p][0];
      const double p0x = C.tx[t0], p0y = C.ty[t0], p0z = C.tz[t0];
      double avx = 0., avy = 0., avz = 0., acx = 0., acy = 0., acz = 0., asum = 0.;
      for (int k = 1; k + 1 < m; ++k) {
        const int ta_ = C.poly[p][k], tb_ = C.poly[p][k + 1];
        const double ax = C.tx[ta_] - p0x, ay = C.ty[ta_] - p0y, az = C.tz[ta_] - p0z;
        const double bx = C.tx[tb_] - p0x, by = C.ty[tb_] - p0y, bz = C.tz[tb_] - p0z;
        const double crx = ay * bz - az * by, cry = az * bx - ax * bz, crz = ax * by - ay * bx;
        const double a = 0.5 * sqrt(v3dot(crx, cry, crz, crx, cry, crz));
        avx = avx + 0.5 * crx;
        avy = avy + 0.5 * cry;
        avz = avz + 0.5 * crz;
        asum += a;
        const double s3 = a / 3.0;
        acx = acx + s3 * ((p0x + C.tx[ta_]) + C.tx[tb_]);
        acy = acy + s3 * ((p0y + C.ty[ta_]) + C.ty[tb_]);
        acz = acz + s3 * ((p0z + C.tz[ta_]) + C.tz[tb_]);
      }
      area = sqrt(v3dot(avx, avy, avz, avx, avy, avz));
      if (asum > 0) {
        const double inv = 1.0 / asum;
        cx = inv * acx;
        cy = inv * acy;
        cz = inv * acz;
      } else {
        cx = p0x;
        cy = p0y;
        cz = p0z;
      }
    }
    C.farea[p] = area;
    C.fcx[p] = cx;
    C.fcy[p] = cy;
    C.fcz[p] = cz;
  }
  if (__ballot(over)) return 3;
  wave_sync();
  // a periodic cell must not touch the initial region (a wall face of positive area)
  bool wall = false;
  for (int p = lane; p < np; p += 64) wall = wall || (C.pnb[p] < 0 && C.farea[p] > 0.);
  if (D.periodic && __ballot(wall)) return 1;
  if (lane != 0) return 0;
  // lane 0: the cell's volume and centroid over the planes and their fans in order, then the faces
  double vol = 0., cenx = 0., ceny = 0., cenz = 0.;
  int cnt = 0;
  for (int p = 0; p < np; ++p) {
    const int m = C.pn[p];
    if (m < 3) continue;
    const int t0 = C.poly[p][0];
    const double p0x = C.tx[t0], p0y = C.ty[t0], p0z = C.tz[t0];
    for (int k = 1; k + 1 < m; ++k) {
      const int ta_ = C.poly[p][k], tb_ = C.poly[p][k + 1];
      const double ax = C.tx[ta_], ay = C.ty[ta_], az = C.tz[ta_], bx = C.tx[tb_], by = C.ty[tb_], bz = C.tz[tb_];
      const double crx = ay * bz - az * by, cry = az * bx - ax * bz, crz = ax * by - ay * bx;
      const double tv = fabs(v3dot(p0x, p0y, p0z, crx, cry, crz)) / 6.0;
      vol += tv;
      const double s4 = tv / 4.0;
      cenx = cenx + s4 * ((p0x + ax) + bx);
      ceny = ceny + s4 * ((p0y + ay) + by);
      cenz = cenz + s4 * ((p0z + az) + bz);
    }
    if (C.farea[p] > 0. && v2_emits(B, D.cell_base, C.pnb[p], i)) ++cnt;
  }
  const int64_t cb = D.cell_base, N = B.C.n;
  const int32_t start = atomicAdd(&E.nf_tmp, cnt);
  if ((int64_t)start + cnt > D.face_cap) {
    V.femit[cb + s + i] = 0;
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
    return 0;
  }
  V.vst[gi] = start;
  V.femit[cb + s + i] = cnt;
  V2Face* out = V.vf + D.face_base + start;
  int q = 0;
  for (int p = 0; p < np; ++p) {
    if (!(C.farea[p] > 0.) || !v2_emits(B, cb, C.pnb[p], i)) continue;
    const double inv = 1.0 / sqrt(v3dot(C.px[p], C.py[p], C.pz[p], C.px[p], C.py[p], C.pz[p]));
    V2Face f;
    f.area = C.farea[p];
    f.cx = C.fcx[p];
    f.cy = C.fcy[p];
    f.cz = C.fcz[p];
    f.nx = inv * C.px[p];
    f.ny = inv * C.py[p];
    f.nz = inv * C.pz[p];
    f.nb = C.pnb[p] < 0 ? -1 : C.pnb[p];
    f.code = C.pcode[p];
    out[q++] = f;
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

// grid: HFV_BLOCK one-wavefront blocks per cell chunk, one cell each.  WIDE: only the cells that the first
// pass marked.
template <bool WIDE>
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
  if (WIDE && V.femit[cb + s + i] != HFV_V2_WIDE) return;
  const double xi = B.C.loc[gi], yi = B.C.loc[N + gi], zi = B.C.loc[2 * N + gi];
  const int kmax = max(6, max(D.vnb[0], max(D.vnb[1], D.vnb[2])));
  int np = 0, nt = 0, r = 1;
  double r2 = 0.;
  for (int K = 2; K <= kmax && r == 1; ++K) {
    r = v3_build<!WIDE>(B, D, V, s, i, xi, yi, zi, K, C, np, nt, r2);
    if (r == 0) r = v3_finish(B, V, D, E, s, gi, i, xi, yi, zi, C, np, nt, r2);
  }
  if (__lane_id() != 0) return;
  if (r == 2 && !WIDE) {
    V.femit[cb + s + i] = HFV_V2_WIDE;
  } else if (r != 0) {
    V.femit[cb + s + i] = 0;
    atomicCAS(&B.cnt->err, 0, E_MESH);
    atomicCAS(&E.err, 0, E_MESH);
  }
}

}  // namespace hfv
