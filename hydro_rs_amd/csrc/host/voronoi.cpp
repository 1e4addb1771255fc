// Host Voronoi tessellation for the finite-volume path.
//
// Replaces the external crate meshless_voronoi 0.6.1 (features rayon, rug; Cargo.lock:434-437), which
// the reference calls at src/space.rs:238-262 and :574-587.  The crate is not vendored, so this is an
// own implementation of the same "meshless" idea: every generator's cell is built independently by
// clipping an initial convex region with the bisector half-spaces of its nearest neighbours, visited
// in order of distance, until the security-radius criterion (no unvisited neighbour can cut the cell)
// holds.  Faces are emitted once, by the lower-index cell of each pair; the right cell of a face is
// described by its index and the periodic shift of its image.  Output layout (host arrays):
//   faces:  area[F], normal[3F] (unit, left -> right image), centroid[3F] (left cell's frame),
//           left[F], right[F] (-1 for a wall face), shift[3F] (image offset of the right cell)
//   cells:  volume[N], centroid[3N], search_radius[N], offset[N], count[N]
//   csr:    conn[sum(count)], each cell's faces in ascending face index.
// 1D follows the semantics pinned by the reference test test_init_1d (src/space.rs:1453-1491).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../../include/hfv_host.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// A clip removes the vertices beyond a plane by more than this fraction of its offset d: vertices that lie
// on it (exact lattices: several bisectors through one vertex) test as beyond by a rounding error, and
// cutting them off would leave slivers of zero width.  The device rebuild uses the same threshold.
constexpr double kClipTol = 1e-12;

// A half-space n.y <= d in coordinates relative to the generator.
struct Plane {
  V3 n;
  double d;
  int64_t nb;     // neighbour index, or -1 - wall for box walls
  int8_t sc[3];   // periodic image code of the neighbour (-1, 0, +1 per axis)
};

struct FaceOut {
  int64_t nb;
  int8_t sc[3];
  double area;
  V3 centroid;  // relative to generator
  V3 normal;
};

struct CellOut {
  double volume;
  V3 centroid;  // relative
  double rmax;
  std::vector<FaceOut> faces;
};

// ----------------------------------------------------------------------------- 3D
// Convex cell stored by its dual: every vertex is a triangle of three plane ids, oriented so that
// det(n_a, n_b, n_c) > 0 for the initial cube; clipping keeps the orientation consistent.
struct Tri {
  int p[3];
  V3 pos;
};

inline bool intersect3(const Plane& a, const Plane& b, const Plane& c, V3* out) {
  V3 bc = cross(b.n, c.n);
  double det = dot(a.n, bc);
  if (det == 0.0 || !std::isfinite(det)) return false;
  V3 r = (a.d * bc) + (b.d * cross(c.n, a.n)) + (c.d * cross(a.n, b.n));
  *out = (1.0 / det) * r;
  return true;
}

struct Cell3 {
  std::vector<Plane> planes;
  std::vector<Tri> tris;
  std::vector<Tri> scratch;
  std::vector<int> removed;

  void init_box(V3 lo, V3 hi) {  // lo/hi relative to generator; lo < 0 < hi
    planes.clear();
    tris.clear();
    // walls: 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z
    planes.push_back({{-1, 0, 0}, -lo.x, -1, {0, 0, 0}});
    planes.push_back({{1, 0, 0}, hi.x, -2, {0, 0, 0}});
    planes.push_back({{0, -1, 0}, -lo.y, -3, {0, 0, 0}});
    planes.push_back({{0, 1, 0}, hi.y, -4, {0, 0, 0}});
    planes.push_back({{0, 0, -1}, -lo.z, -5, {0, 0, 0}});
    planes.push_back({{0, 0, 1}, hi.z, -6, {0, 0, 0}});
    for (int sx = 0; sx < 2; ++sx)
      for (int sy = 0; sy < 2; ++sy)
        for (int sz = 0; sz < 2; ++sz) {
          int a = sx, b = 2 + sy, c = 4 + sz;
          double det = dot(planes[a].n, cross(planes[b].n, planes[c].n));
          Tri t;
          if (det > 0) {
            t.p[0] = a; t.p[1] = b; t.p[2] = c;
          } else {
            t.p[0] = b; t.p[1] = a; t.p[2] = c;
          }
          intersect3(planes[t.p[0]], planes[t.p[1]], planes[t.p[2]], &t.pos);
          tris.push_back(t);
        }
  }

  double rmax2() const {
    double r = 0;
    for (const Tri& t : tris) r = std::max(r, dot(t.pos, t.pos));
    return r;
  }

  // Where three planes meet in no point (their normals coplanar, as where more than three bisectors pass
  // through one vertex of an exact lattice): the point where the edge (x, y) from the removed vertex t to
  // the kept vertex across it (the triangle with the edge (y, x)) crosses P; t's own position if there is
  // none.
  V3 edge_crossing(const Plane& P, const Tri& t, int x, int y) const {
    for (const Tri& u : tris) {
      bool has = false;
      for (int k = 0; k < 3; ++k) has = has || (u.p[k] == y && u.p[(k + 1) % 3] == x);
      if (!has) continue;
      const double a = dot(P.n, u.pos) - P.d, b = dot(P.n, t.pos) - P.d;
      if (!(a <= 0.0) || !(b > a)) return t.pos;  // (not a kept vertex against a removed one)
      const double s = a / (a - b);
      return u.pos + (s * (t.pos - u.pos));
    }
    return t.pos;
  }

  // Returns false if the plane does not cut the cell.
  bool clip(const Plane& P) {
    removed.clear();
    for (int i = 0; i < (int)tris.size(); ++i)
      if (dot(P.n, tris[i].pos) > P.d + kClipTol * P.d) removed.push_back(i);
    if (removed.empty()) return false;
    if (removed.size() == tris.size()) return false;  // cannot happen for a bisector of a neighbour
    int pid = (int)planes.size();
    planes.push_back(P);
    scratch.clear();
    // boundary edges of the removed region -> new triangles
    for (int ri : removed) {
      const Tri& t = tris[ri];
      for (int e = 0; e < 3; ++e) {
        int x = t.p[e], y = t.p[(e + 1) % 3];
        bool internal = false;
        for (int rj : removed) {
          const Tri& u = tris[rj];
          for (int k = 0; k < 3; ++k)
            if (u.p[k] == y && u.p[(k + 1) % 3] == x) {
              internal = true;
              break;
            }
          if (internal) break;
        }
        if (!internal) {
          Tri nt;
          nt.p[0] = x;
          nt.p[1] = y;
          nt.p[2] = pid;
          if (!intersect3(planes[x], planes[y], P, &nt.pos)) nt.pos = edge_crossing(P, t, x, y);
          scratch.push_back(nt);
        }
      }
    }
    // compact: drop removed, append new
    std::vector<char> dead(tris.size(), 0);
    for (int ri : removed) dead[ri] = 1;
    size_t w = 0;
    for (size_t i = 0; i < tris.size(); ++i)
      if (!dead[i]) tris[w++] = tris[i];
    tris.resize(w);
    for (const Tri& t : scratch) tris.push_back(t);
    return true;
  }

  void finish(CellOut* out) {
    out->faces.clear();
    size_t np = planes.size();
    // for every plane, list its triangles rotated to (plane, u, v)
    std::vector<std::vector<std::pair<int, int>>> fan(np);
    std::vector<std::vector<int>> fan_tri(np);
    for (int ti = 0; ti < (int)tris.size(); ++ti) {
      const Tri& t = tris[ti];
      for (int k = 0; k < 3; ++k) {
        fan[t.p[k]].push_back({t.p[(k + 1) % 3], t.p[(k + 2) % 3]});
        fan_tri[t.p[k]].push_back(ti);
      }
    }
    double vol = 0;
    V3 cen = {0, 0, 0};
    std::vector<V3> poly;
    for (size_t p = 0; p < np; ++p) {
      const auto& f = fan[p];
      if (f.size() < 3) continue;
      // chain: from (u, v) the next triangle around p has u' = v
      poly.clear();
      int cur = 0;
      for (size_t step = 0; step < f.size(); ++step) {
        poly.push_back(tris[fan_tri[p][cur]].pos);
        int want = f[cur].second;
        int nxt = -1;
        for (size_t q = 0; q < f.size(); ++q)
          if (f[q].first == want) {
            nxt = (int)q;
            break;
          }
        if (nxt < 0) break;
        cur = nxt;
        if (cur == 0) break;
      }
      if (poly.size() < 3) continue;
      V3 p0 = poly[0];
      V3 avec = {0, 0, 0};
      V3 acen = {0, 0, 0};
      double asum = 0;
      for (size_t k = 1; k + 1 < poly.size(); ++k) {
        V3 c = cross(poly[k] - p0, poly[k + 1] - p0);
        double a = 0.5 * norm(c);
        avec = avec + (0.5 * c);
        asum += a;
        acen = acen + ((a / 3.0) * (p0 + poly[k] + poly[k + 1]));
        // tetrahedron with the generator (origin)
        double tv = std::fabs(dot(p0, cross(poly[k], poly[k + 1]))) / 6.0;
        vol += tv;
        cen = cen + ((tv / 4.0) * (p0 + poly[k] + poly[k + 1]));
      }
      double area = norm(avec);
      if (!(area > 0)) continue;
      FaceOut fo;
      fo.nb = planes[p].nb;
      std::memcpy(fo.sc, planes[p].sc, 3);
      fo.area = area;
      fo.centroid = (asum > 0) ? (1.0 / asum) * acen : p0;
      V3 n = planes[p].n;
      fo.normal = (1.0 / norm(n)) * n;
      out->faces.push_back(fo);
    }
    out->volume = vol;
    out->centroid = vol > 0 ? (1.0 / vol) * cen : V3{0, 0, 0};
    out->rmax = std::sqrt(rmax2());
  }
};

// ----------------------------------------------------------------------------- 2D
// Convex polygon as a cyclic list of lines; vertex k is the intersection of lines k and k+1.
struct Cell2 {
  std::vector<Plane> lines;
  std::vector<V3> verts;
  std::vector<Plane> nl;

  static bool isect(const Plane& a, const Plane& b, V3* out) {
    double det = a.n.x * b.n.y - a.n.y * b.n.x;
    if (det == 0.0) return false;
    out->x = (a.d * b.n.y - b.d * a.n.y) / det;
    out->y = (a.n.x * b.d - b.n.x * a.d) / det;
    out->z = 0;
    return true;
  }
  void recompute() {
    size_t m = lines.size();
    verts.resize(m);
    for (size_t k = 0; k < m; ++k) {
      if (!isect(lines[k], lines[(k + 1) % m], &verts[k])) verts[k] = {0, 0, 0};
    }
  }
  void init_box(V3 lo, V3 hi) {
    lines.clear();
    // counter-clockwise: bottom (-y), right (+x), top (+y), left (-x)
    lines.push_back({{0, -1, 0}, -lo.y, -3, {0, 0, 0}});
    lines.push_back({{1, 0, 0}, hi.x, -2, {0, 0, 0}});
    lines.push_back({{0, 1, 0}, hi.y, -4, {0, 0, 0}});
    lines.push_back({{-1, 0, 0}, -lo.x, -1, {0, 0, 0}});
    recompute();
  }
  double rmax2() const {
    double r = 0;
    for (const V3& v : verts) r = std::max(r, v.x * v.x + v.y * v.y);
    return r;
  }
  bool clip(const Plane& P) {
    size_t m = lines.size();
    std::vector<char> out(m);
    int nout = 0;
    for (size_t k = 0; k < m; ++k) {
      out[k] = (P.n.x * verts[k].x + P.n.y * verts[k].y) > P.d + kClipTol * P.d;
      nout += out[k];
    }
    if (nout == 0 || nout == (int)m) return false;
    // find first outside vertex after an inside one
    size_t start = 0;
    for (size_t k = 0; k < m; ++k)
      if (out[k] && !out[(k + m - 1) % m]) {
        start = k;
        break;
      }
    // vertices start .. start+nout-1 are outside (contiguous by convexity). Vertex k sits between
    // lines k and k+1, so lines start+1 .. start+nout-1 disappear; the new line goes after line start.
    nl.clear();
    size_t keep_from = (start + nout) % m;  // line index after the removed run
    nl.push_back(P);
    // lines from keep_from ... start (inclusive), cyclic
    size_t k = keep_from;
    while (true) {
      nl.push_back(lines[k]);
      if (k == start) break;
      k = (k + 1) % m;
    }
    // nl = [P, keep_from, ..., start]; rotate so order stays ccw: ..., start, P, keep_from, ...
    std::rotate(nl.begin(), nl.begin() + 1, nl.end());
    lines.swap(nl);
    recompute();
    return true;
  }
  void finish(CellOut* out) {
    out->faces.clear();
    size_t m = lines.size();
    double area = 0;
    V3 c = {0, 0, 0};
    for (size_t k = 0; k < m; ++k) {
      // line k spans vertices k-1 .. k
      V3 a = verts[(k + m - 1) % m], b = verts[k];
      double cr = a.x * b.y - b.x * a.y;
      area += 0.5 * cr;
      c = c + ((cr / 6.0) * (a + b));
      double len = std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
      if (!(len > 0)) continue;
      FaceOut fo;
      fo.nb = lines[k].nb;
      std::memcpy(fo.sc, lines[k].sc, 3);
      fo.area = len;
      fo.centroid = 0.5 * (a + b);
      V3 n = lines[k].n;
      fo.normal = (1.0 / std::sqrt(n.x * n.x + n.y * n.y)) * n;
      fo.normal.z = 0;
      out->faces.push_back(fo);
    }
    out->volume = area;
    out->centroid = area > 0 ? (1.0 / area) * c : V3{0, 0, 0};
    out->rmax = std::sqrt(rmax2());
  }
};

struct Grid {
  int dim;
  double box[3];
  int periodic;
  int nb[3];
  double w[3];
  std::vector<int64_t> start;  // bin -> first point (sorted by bin)
  std::vector<int64_t> order;

  int bin_of(double x, int a) const {
    int b = (int)std::floor(x / w[a]);
    if (b < 0) b = 0;
    if (b >= nb[a]) b = nb[a] - 1;
    return b;
  }
  void build(int64_t n, const double* loc) {
    double vol = 1;
    for (int a = 0; a < dim; ++a) vol *= box[a];
    // mean generator spacing (sqrt / cbrt, not pow: the device rebuild evaluates the same correctly
    // rounded or libm call, so both grids agree bin for bin)
    const double per_gen = vol / (double)std::max<int64_t>(n, 1);
    double spacing = dim == 1 ? per_gen : dim == 2 ? std::sqrt(per_gen) : std::cbrt(per_gen);
    for (int a = 0; a < 3; ++a) {
      if (a < dim) {
        nb[a] = std::max(1, (int)std::floor(box[a] / spacing));
        w[a] = box[a] / nb[a];
      } else {
        nb[a] = 1;
        w[a] = 1;
      }
    }
    int64_t nbins = (int64_t)nb[0] * nb[1] * nb[2];
    std::vector<int64_t> cnt(nbins + 1, 0);
    std::vector<int64_t> key(n);
    for (int64_t i = 0; i < n; ++i) {
      int b[3] = {0, 0, 0};
      for (int a = 0; a < dim; ++a) b[a] = bin_of(loc[3 * i + a], a);
      key[i] = ((int64_t)b[0] * nb[1] + b[1]) * nb[2] + b[2];
      cnt[key[i] + 1]++;
    }
    for (int64_t b = 0; b < nbins; ++b) cnt[b + 1] += cnt[b];
    start = cnt;
    order.resize(n);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; ++i) order[fill[key[i]]++] = i;
  }
};

struct Cand {
  double d2;
  int64_t j;
  int8_t sc[3];
  V3 rel;
};

// Build one cell. Returns 0 on success, 1 if the candidate block was too small.
template <class CellT>
int build_cell(const Grid& g, int64_t n, const double* loc, int64_t i, int K, CellT& cell, std::vector<Cand>& cands,
               CellOut* out) {
  int dim = g.dim;
  V3 xi = {loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]};
  int bi[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) bi[a] = g.bin_of((&xi.x)[a], a);
  cands.clear();
  int kr[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) kr[a] = K;
  for (int dx = -kr[0]; dx <= kr[0]; ++dx)
    for (int dy = -kr[1]; dy <= kr[1]; ++dy)
      for (int dz = -kr[2]; dz <= kr[2]; ++dz) {
        int b[3] = {bi[0] + dx, bi[1] + dy, bi[2] + dz};
        int8_t sc[3] = {0, 0, 0};
        bool skip = false;
        for (int a = 0; a < dim; ++a) {
          if (b[a] < 0 || b[a] >= g.nb[a]) {
            if (!g.periodic) {
              skip = true;
              break;
            }
            // wrap once (the block is narrower than the box)
            if (b[a] < 0) {
              b[a] += g.nb[a];
              sc[a] = -1;
            } else {
              b[a] -= g.nb[a];
              sc[a] = 1;
            }
            if (b[a] < 0 || b[a] >= g.nb[a]) {
              skip = true;
              break;
            }
          }
        }
        if (skip) continue;
        int64_t key = ((int64_t)b[0] * g.nb[1] + b[1]) * g.nb[2] + b[2];
        for (int64_t q = g.start[key]; q < g.start[key + 1]; ++q) {
          int64_t j = g.order[q];
          V3 xj = {loc[3 * j] + sc[0] * g.box[0], loc[3 * j + 1] + sc[1] * g.box[1], loc[3 * j + 2] + sc[2] * g.box[2]};
          if (j == i && sc[0] == 0 && sc[1] == 0 && sc[2] == 0) continue;
          V3 r = xj - xi;
          Cand c;
          c.d2 = dot(r, r);
          c.j = j;
          std::memcpy(c.sc, sc, 3);
          c.rel = r;
          cands.push_back(c);
        }
      }
  std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
    if (a.d2 != b.d2) return a.d2 < b.d2;
    if (a.j != b.j) return a.j < b.j;
    return std::memcmp(a.sc, b.sc, 3) < 0;
  });
  // initial region
  V3 lo, hi;
  double reach = 0;
  for (int a = 0; a < dim; ++a) reach = std::max(reach, K * g.w[a]);
  if (g.periodic) {
    lo = {-reach, -reach, -reach};
    hi = {reach, reach, reach};
  } else {
    lo = {-xi.x, -xi.y, -xi.z};
    hi = {g.box[0] - xi.x, g.box[1] - xi.y, g.box[2] - xi.z};
  }
  if (dim == 2) {
    lo.z = -1;
    hi.z = 1;
  }
  cell.init_box(lo, hi);
  double r2 = cell.rmax2();
  for (const Cand& c : cands) {
    if (0.25 * c.d2 > r2) break;  // security radius: the bisector lies beyond every vertex
    Plane P;
    P.n = c.rel;
    P.d = 0.5 * c.d2;
    P.nb = c.j;
    std::memcpy(P.sc, c.sc, 3);
    if (cell.clip(P)) r2 = cell.rmax2();
  }
  // completeness: every unvisited point lies farther than K * w from the generator
  double minw = 1e300;
  for (int a = 0; a < dim; ++a) minw = std::min(minw, K * g.w[a]);
  if (4.0 * r2 > minw * minw) return 1;
  cell.finish(out);
  if (g.periodic)
    for (const FaceOut& f : out->faces)
      if (f.nb < 0) return 1;
  return 0;
}

}  // namespace

struct hfv_mesh {
  int dim;
  int64_t n, nf;
  std::vector<double> area, normal, centroid, shift;
  std::vector<int64_t> left, right;
  std::vector<double> volume, cell_centroid, search_radius;
  std::vector<int64_t> offset, count, conn;
};

static void build_csr(hfv_mesh* m) {
  int64_t n = m->n, nf = m->nf;
  m->count.assign(n, 0);
  for (int64_t f = 0; f < nf; ++f) {
    m->count[m->left[f]]++;
    if (m->right[f] >= 0) m->count[m->right[f]]++;
  }
  m->offset.assign(n, 0);
  int64_t acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    m->offset[i] = acc;
    acc += m->count[i];
  }
  m->conn.assign(acc, 0);
  std::vector<int64_t> fill(m->offset);
  for (int64_t f = 0; f < nf; ++f) {
    m->conn[fill[m->left[f]]++] = f;
    if (m->right[f] >= 0) m->conn[fill[m->right[f]]++] = f;
  }
}

static int build_1d(hfv_mesh* m, const double* loc, double L, int periodic) {
  int64_t n = m->n;
  std::vector<double> x(n);
  for (int64_t i = 0; i < n; ++i) x[i] = loc[3 * i];
  for (int64_t i = 0; i + 1 < n; ++i)
    if (!(x[i] < x[i + 1])) return HFV_ERR_INVALID;  // generators must be sorted along x
  int64_t nf = periodic ? n : n + 1;
  m->nf = nf;
  m->area.assign(nf, 1.0);
  m->normal.assign(3 * nf, 0.0);
  m->centroid.assign(3 * nf, 0.0);
  m->shift.assign(3 * nf, 0.0);
  m->left.assign(nf, 0);
  m->right.assign(nf, 0);
  for (int64_t f = 0; f < nf; ++f) m->normal[3 * f] = 1.0;
  for (int64_t k = 1; k < n; ++k) {
    m->centroid[3 * k] = 0.5 * (x[k - 1] + x[k]);
    m->left[k] = k - 1;
    m->right[k] = k;
  }
  if (!periodic) {
    m->normal[0] = -1.0;
    m->left[0] = 0;
    m->right[0] = -1;
    m->centroid[3 * n] = L;
    m->left[n] = n - 1;
    m->right[n] = -1;
  } else {
    m->centroid[0] = 0.5 * (x[n - 1] + (x[0] + L));
    m->left[0] = n - 1;
    m->right[0] = 0;
    m->shift[0] = L;
  }
  m->volume.assign(n, 0);
  m->cell_centroid.assign(3 * n, 0);
  m->search_radius.assign(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    double xl, xr;
    if (!periodic) {
      xl = m->centroid[3 * i];
      xr = m->centroid[3 * (i + 1)];
    } else {
      xl = (i == 0) ? m->centroid[0] - L : m->centroid[3 * i];
      xr = (i == n - 1) ? m->centroid[0] : m->centroid[3 * (i + 1)];
    }
    m->volume[i] = xr - xl;
    m->cell_centroid[3 * i] = 0.5 * (xl + xr);
    m->search_radius[i] = 1.5 * m->volume[i];
  }
  build_csr(m);
  return HFV_OK;
}

extern "C" int hfv_voronoi_build(int dim, const double* box, int periodic, int64_t n, const double* loc, int n_threads,
                                 hfv_mesh** out) {
  return hfv_voronoi_build_partial(dim, box, periodic, n, loc, nullptr, n_threads, out);
}

extern "C" int hfv_voronoi_build_partial(int dim, const double* box, int periodic, int64_t n, const double* loc,
                                         const uint8_t* mask, int n_threads, hfv_mesh** out) {
  if (!out || !loc || !box || n <= 0 || dim < 1 || dim > 3) return HFV_ERR_INVALID;
  hfv_mesh* m = new hfv_mesh();
  m->dim = dim;
  m->n = n;
  if (dim == 1) {
    int e = build_1d(m, loc, box[0], periodic);
    if (e) {
      delete m;
      return e;
    }
    *out = m;
    return HFV_OK;
  }
  Grid g;
  g.dim = dim;
  g.periodic = periodic;
  for (int a = 0; a < 3; ++a) g.box[a] = box[a];
  g.build(n, loc);
  std::vector<CellOut> cells(n);
  // candidate blocks of K = 2, 3, .. bins around a generator, up to the whole grid (cells next to large
  // voids need wide blocks)
  const int kmax = std::max(6, std::max(g.nb[0], std::max(g.nb[1], g.nb[2])));
  std::atomic<int> failed(0);
#ifdef _OPENMP
  int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
  {
    Cell3 c3;
    Cell2 c2;
    std::vector<Cand> cands;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < n; ++i) {
      if (mask && !mask[i]) continue;  // a partial build leaves unmasked cells unbuilt
      int ok = 1;
      for (int K = 2; K <= kmax; ++K) {
        int r = (dim == 3) ? build_cell(g, n, loc, i, K, c3, cands, &cells[i])
                           : build_cell(g, n, loc, i, K, c2, cands, &cells[i]);
        if (r == 0) {
          ok = 0;
          break;
        }
      }
      if (ok) failed.store(1);
    }
  }
  if (failed.load()) {
    delete m;
    return HFV_ERR_MESH;
  }
  // faces: emitted by the lower-index cell of each pair of built cells, by the built cell of a pair with
  // an unbuilt one (partial builds), and walls by their cell
  auto emits = [&](const FaceOut& f, int64_t i) { return f.nb < 0 || f.nb > i || (mask && !mask[f.nb]); };
  std::vector<int64_t> nfc(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = 0;
    for (const FaceOut& f : cells[i].faces)
      if (emits(f, i)) c++;
    nfc[i + 1] = c;
  }
  for (int64_t i = 0; i < n; ++i) nfc[i + 1] += nfc[i];
  int64_t nf = nfc[n];
  m->nf = nf;
  m->area.assign(nf, 0);
  m->normal.assign(3 * nf, 0);
  m->centroid.assign(3 * nf, 0);
  m->shift.assign(3 * nf, 0);
  m->left.assign(nf, 0);
  m->right.assign(nf, 0);
  m->volume.assign(n, 0);
  m->cell_centroid.assign(3 * n, 0);
  m->search_radius.assign(n, 0);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) {
    const CellOut& c = cells[i];
    const double* xi = &loc[3 * i];
    m->volume[i] = c.volume;
    m->cell_centroid[3 * i + 0] = xi[0] + c.centroid.x;
    m->cell_centroid[3 * i + 1] = xi[1] + c.centroid.y;
    m->cell_centroid[3 * i + 2] = dim == 3 ? xi[2] + c.centroid.z : xi[2];
    m->search_radius[i] = 1.5 * 2.0 * c.rmax;
    int64_t f = nfc[i];
    for (const FaceOut& fo : c.faces) {
      if (!emits(fo, i)) continue;
      m->area[f] = fo.area;
      m->normal[3 * f + 0] = fo.normal.x;
      m->normal[3 * f + 1] = fo.normal.y;
      m->normal[3 * f + 2] = dim == 3 ? fo.normal.z : 0.0;
      m->centroid[3 * f + 0] = xi[0] + fo.centroid.x;
      m->centroid[3 * f + 1] = xi[1] + fo.centroid.y;
      m->centroid[3 * f + 2] = dim == 3 ? xi[2] + fo.centroid.z : xi[2];
      m->left[f] = i;
      m->right[f] = fo.nb < 0 ? -1 : fo.nb;
      for (int a = 0; a < 3; ++a) m->shift[3 * f + a] = (double)fo.sc[a] * box[a];
      f++;
    }
  }
  build_csr(m);
  *out = m;
  return HFV_OK;
}

extern "C" void hfv_mesh_free(hfv_mesh* m) { delete m; }

extern "C" int hfv_mesh_view(const hfv_mesh* m, hfv_mesh_arrays* v) {
  if (!m || !v) return HFV_ERR_INVALID;
  v->n_cells = m->n;
  v->n_faces = m->nf;
  v->n_conn = (int64_t)m->conn.size();
  v->area = m->area.data();
  v->normal = m->normal.data();
  v->centroid = m->centroid.data();
  v->shift = m->shift.data();
  v->left = m->left.data();
  v->right = m->right.data();
  v->volume = m->volume.data();
  v->cell_centroid = m->cell_centroid.data();
  v->search_radius = m->search_radius.data();
  v->offset = m->offset.data();
  v->count = m->count.data();
  v->conn = m->conn.data();
  return HFV_OK;
}
