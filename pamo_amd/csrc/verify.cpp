// EXACT MESH VERIFIER (libpamo_verify.so; CLI: python -m pamo_amd.verify).
//
// Checks a triangle mesh for (1) manifoldness and watertightness,
// (2) degenerate faces and (3) self-intersections, with exact predicates.
// It replaces the reference's offline CGAL EPICK checker
// (simp_cuda/safe_project/scripts/cgal_intersect/main.cpp:12,66), which is
// not buildable here (no CGAL/GMP).
//
// Predicates: orient3d / orient2d on float coordinates, evaluated in double
// with Shewchuk's static error bound and, when that filter cannot decide, an
// exact expansion-arithmetic evaluation (J. R. Shewchuk, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", 1997:
// Two-Sum, Two-Diff, Dekker split / Two-Product, Scale-Expansion and
// Expansion-Sum with zero elimination).  Compile with -ffp-contract=off.
//
// Intersection semantics (closed triangles, as CGAL's do_intersect):
//   * no shared vertex index: any common point is an intersection;
//   * one shared vertex p: an intersection iff the triangles share a point
//     other than p;
//   * one shared edge: an intersection iff the triangles are coplanar and lie
//     on the same side of the edge (they fold over each other);
//   * three shared indices (duplicate face): an intersection.
// Degenerate (collinear or repeated-index) faces are counted separately and
// skipped by the intersection test.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace vf {

// ------------------------------------------------ expansion arithmetic
typedef std::vector<double> Exp;

static inline void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  double bv = x - a, av = x - bv, br = b - bv, ar = a - av;
  y = ar + br;
}
static inline void two_diff(double a, double b, double& x, double& y) {
  x = a - b;
  double bv = a - x, av = x + bv, br = bv - b, ar = a - av;
  y = ar + br;
}
static const double SPLITTER = 134217729.0;  // 2^27 + 1
static inline void split(double a, double& hi, double& lo) {
  double c = SPLITTER * a, abig = c - a;
  hi = c - abig;
  lo = a - hi;
}
static inline void two_product(double a, double b, double& x, double& y) {
  x = a * b;
  double ahi, alo, bhi, blo;
  split(a, ahi, alo);
  split(b, bhi, blo);
  double err1 = x - ahi * bhi, err2 = err1 - alo * bhi, err3 = err2 - ahi * blo;
  y = alo * blo - err3;
}
// sum of two nonoverlapping expansions (increasing magnitude), zero-eliminated
static Exp exp_sum(const Exp& e, const Exp& f) {
  Exp h;
  h.reserve(e.size() + f.size());
  // grow-expansion based linear sum (Shewchuk's Linear-Expansion-Sum is
  // replaced by repeated Grow-Expansion, which is simpler and still exact)
  Exp cur = e;
  for (double b : f) {
    Exp nxt;
    nxt.reserve(cur.size() + 1);
    double q = b;
    for (double ei : cur) {
      double x, y;
      two_sum(q, ei, x, y);
      if (y != 0.0) nxt.push_back(y);
      q = x;
    }
    if (q != 0.0 || nxt.empty()) nxt.push_back(q);
    cur.swap(nxt);
  }
  return cur;
}
static Exp exp_scale(const Exp& e, double b) {
  Exp h;
  if (e.empty()) return h;
  double q, y;
  two_product(e[0], b, q, y);
  if (y != 0.0) h.push_back(y);
  for (size_t i = 1; i < e.size(); i++) {
    double p1, p0, s, t;
    two_product(e[i], b, p1, p0);
    two_sum(q, p0, s, t);
    if (t != 0.0) h.push_back(t);
    two_sum(p1, s, q, t);
    if (t != 0.0) h.push_back(t);
  }
  if (q != 0.0 || h.empty()) h.push_back(q);
  return h;
}
static Exp exp_mul(const Exp& e, const Exp& f) {
  Exp acc;
  for (double b : f) acc = exp_sum(acc, exp_scale(e, b));
  return acc;
}
static Exp exp_neg(Exp e) { for (double& x : e) x = -x; return e; }
static int exp_sign(const Exp& e) {
  for (size_t i = e.size(); i-- > 0;) if (e[i] != 0.0) return e[i] > 0 ? 1 : -1;
  return 0;
}
static Exp diff2(double a, double b) { double x, y; two_diff(a, b, x, y); Exp e; if (y != 0.0) e.push_back(y); e.push_back(x); return e; }

static const double EPS = 1.1102230246251565e-16;  // 2^-53
static const double O3D_BOUND = (7.0 + 56.0 * EPS) * EPS;
static const double CCW_BOUND = (3.0 + 16.0 * EPS) * EPS;

struct P3 { double x, y, z; };

static int orient3d_exact(const P3& a, const P3& b, const P3& c, const P3& d) {
  Exp adx = diff2(a.x, d.x), ady = diff2(a.y, d.y), adz = diff2(a.z, d.z);
  Exp bdx = diff2(b.x, d.x), bdy = diff2(b.y, d.y), bdz = diff2(b.z, d.z);
  Exp cdx = diff2(c.x, d.x), cdy = diff2(c.y, d.y), cdz = diff2(c.z, d.z);
  Exp t1 = exp_sum(exp_mul(bdx, cdy), exp_neg(exp_mul(bdy, cdx)));
  Exp t2 = exp_sum(exp_mul(cdx, ady), exp_neg(exp_mul(cdy, adx)));
  Exp t3 = exp_sum(exp_mul(adx, bdy), exp_neg(exp_mul(ady, bdx)));
  Exp det = exp_sum(exp_sum(exp_mul(adz, t1), exp_mul(bdz, t2)), exp_mul(cdz, t3));
  return exp_sign(det);
}

// sign of det[a-d; b-d; c-d]
static int orient3d(const P3& a, const P3& b, const P3& c, const P3& d) {
  double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
  double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
  double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
  double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
  double bound = O3D_BOUND * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return orient3d_exact(a, b, c, d);
}

struct P2 { double x, y; };
static int orient2d_exact(const P2& a, const P2& b, const P2& c) {
  Exp acx = diff2(a.x, c.x), acy = diff2(a.y, c.y), bcx = diff2(b.x, c.x), bcy = diff2(b.y, c.y);
  return exp_sign(exp_sum(exp_mul(acx, bcy), exp_neg(exp_mul(acy, bcx))));
}
static int orient2d(const P2& a, const P2& b, const P2& c) {
  double dl = (a.x - c.x) * (b.y - c.y), dr = (a.y - c.y) * (b.x - c.x);
  double det = dl - dr;
  double bound = CCW_BOUND * (fabs(dl) + fabs(dr));
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return orient2d_exact(a, b, c);
}

static inline P2 proj(const P3& p, int drop) {
  if (drop == 0) return {p.y, p.z};
  if (drop == 1) return {p.x, p.z};
  return {p.x, p.y};
}
// axis to drop so that the triangle's projection is non-degenerate; -1 if
// the triangle is degenerate (collinear).
static int proj_axis(const P3& a, const P3& b, const P3& c) {
  for (int drop : {2, 1, 0})
    if (orient2d(proj(a, drop), proj(b, drop), proj(c, drop)) != 0) return drop;
  return -1;
}

// --------------------------------------------------------- 2-D tests
static inline bool on_seg_2d(const P2& p, const P2& q, const P2& r) {  // r collinear with pq
  return std::min(p.x, q.x) <= r.x && r.x <= std::max(p.x, q.x) && std::min(p.y, q.y) <= r.y &&
         r.y <= std::max(p.y, q.y);
}
static bool seg_seg_2d(const P2& p1, const P2& p2, const P2& q1, const P2& q2) {
  int d1 = orient2d(q1, q2, p1), d2 = orient2d(q1, q2, p2), d3 = orient2d(p1, p2, q1), d4 = orient2d(p1, p2, q2);
  if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return true;
  if (d1 == 0 && on_seg_2d(q1, q2, p1)) return true;
  if (d2 == 0 && on_seg_2d(q1, q2, p2)) return true;
  if (d3 == 0 && on_seg_2d(p1, p2, q1)) return true;
  if (d4 == 0 && on_seg_2d(p1, p2, q2)) return true;
  return false;
}
static bool pt_in_tri_2d(const P2& p, const P2& a, const P2& b, const P2& c) {
  int s = orient2d(a, b, c);
  int o1 = orient2d(a, b, p) * s, o2 = orient2d(b, c, p) * s, o3 = orient2d(c, a, p) * s;
  return o1 >= 0 && o2 >= 0 && o3 >= 0;
}
static bool seg_tri_2d(const P2& p, const P2& q, const P2& a, const P2& b, const P2& c) {
  if (pt_in_tri_2d(p, a, b, c) || pt_in_tri_2d(q, a, b, c)) return true;
  return seg_seg_2d(p, q, a, b) || seg_seg_2d(p, q, b, c) || seg_seg_2d(p, q, c, a);
}
static bool tri_tri_2d(const P2 t[3], const P2 u[3]) {
  for (int i = 0; i < 3; i++) if (pt_in_tri_2d(t[i], u[0], u[1], u[2]) || pt_in_tri_2d(u[i], t[0], t[1], t[2])) return true;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (seg_seg_2d(t[i], t[(i + 1) % 3], u[j], u[(j + 1) % 3])) return true;
  return false;
}

// --------------------------------------------------------- 3-D tests
// closed segment pq vs closed triangle abc
static bool seg_tri(const P3& p, const P3& q, const P3& a, const P3& b, const P3& c) {
  int sp = orient3d(a, b, c, p), sq = orient3d(a, b, c, q);
  if ((sp > 0 && sq > 0) || (sp < 0 && sq < 0)) return false;
  if (sp == 0 && sq == 0) {
    int ax = proj_axis(a, b, c);
    if (ax < 0) return false;
    return seg_tri_2d(proj(p, ax), proj(q, ax), proj(a, ax), proj(b, ax), proj(c, ax));
  }
  int s1 = orient3d(p, q, a, b), s2 = orient3d(p, q, b, c), s3 = orient3d(p, q, c, a);
  bool pos = s1 > 0 || s2 > 0 || s3 > 0, neg = s1 < 0 || s2 < 0 || s3 < 0;
  return !(pos && neg);
}

static bool tri_tri_general(const P3 t[3], const P3 u[3]) {
  int o[3], w[3];
  for (int i = 0; i < 3; i++) o[i] = orient3d(t[0], t[1], t[2], u[i]);
  if ((o[0] > 0 && o[1] > 0 && o[2] > 0) || (o[0] < 0 && o[1] < 0 && o[2] < 0)) return false;
  for (int i = 0; i < 3; i++) w[i] = orient3d(u[0], u[1], u[2], t[i]);
  if ((w[0] > 0 && w[1] > 0 && w[2] > 0) || (w[0] < 0 && w[1] < 0 && w[2] < 0)) return false;
  if (o[0] == 0 && o[1] == 0 && o[2] == 0) {
    int ax = proj_axis(t[0], t[1], t[2]);
    if (ax < 0) return false;
    P2 a[3] = {proj(t[0], ax), proj(t[1], ax), proj(t[2], ax)}, b[3] = {proj(u[0], ax), proj(u[1], ax), proj(u[2], ax)};
    return tri_tri_2d(a, b);
  }
  for (int i = 0; i < 3; i++) if (seg_tri(t[i], t[(i + 1) % 3], u[0], u[1], u[2])) return true;
  for (int i = 0; i < 3; i++) if (seg_tri(u[i], u[(i + 1) % 3], t[0], t[1], t[2])) return true;
  return false;
}

// segment p->a (p a vertex of triangle p,c,d) meets the triangle in a point other than p
static bool ray_from_shared(const P3& p, const P3& a, const P3& c, const P3& d) {
  if (orient3d(p, c, d, a) != 0) return false;
  int ax = proj_axis(p, c, d);
  if (ax < 0) return false;
  P2 P = proj(p, ax), A = proj(a, ax), C = proj(c, ax), D = proj(d, ax);
  int s = orient2d(P, C, D);
  return s * orient2d(P, C, A) >= 0 && s * orient2d(P, A, D) >= 0;
}

// triangles (p,a,b) and (p,c,d) sharing exactly vertex p
static bool tri_tri_shared_vertex(const P3& p, const P3& a, const P3& b, const P3& c, const P3& d) {
  if (seg_tri(a, b, p, c, d)) return true;
  if (seg_tri(c, d, p, a, b)) return true;
  if (ray_from_shared(p, a, c, d) || ray_from_shared(p, b, c, d)) return true;
  if (ray_from_shared(p, c, a, b) || ray_from_shared(p, d, a, b)) return true;
  return false;
}

// triangles (p,q,a) and (p,q,b) sharing edge pq
static bool tri_tri_shared_edge(const P3& p, const P3& q, const P3& a, const P3& b) {
  if (orient3d(p, q, a, b) != 0) return false;
  int ax = proj_axis(p, q, a);
  if (ax < 0) return false;
  int sa = orient2d(proj(p, ax), proj(q, ax), proj(a, ax));
  int sb = orient2d(proj(p, ax), proj(q, ax), proj(b, ax));
  return sa * sb >= 0;
}

struct Box { float lo[3], hi[3]; };
static inline bool overlap(const Box& a, const Box& b) {
  for (int k = 0; k < 3; k++) if (a.hi[k] < b.lo[k] || b.hi[k] < a.lo[k]) return false;
  return true;
}

struct BVH {
  struct Node { Box b; int left, right, first, count; };
  std::vector<Node> nodes;
  std::vector<int> items;
  const std::vector<Box>* boxes = nullptr;
  void build(const std::vector<Box>& bx, const std::vector<int>& ids) {
    boxes = &bx; items = ids; nodes.clear();
    if (!items.empty()) rec(0, (int)items.size());
  }
  int rec(int first, int count) {
    Node n; n.left = n.right = -1; n.first = first; n.count = count;
    for (int k = 0; k < 3; k++) { n.b.lo[k] = INFINITY; n.b.hi[k] = -INFINITY; }
    for (int i = first; i < first + count; i++) {
      const Box& x = (*boxes)[items[i]];
      for (int k = 0; k < 3; k++) { n.b.lo[k] = std::min(n.b.lo[k], x.lo[k]); n.b.hi[k] = std::max(n.b.hi[k], x.hi[k]); }
    }
    int idx = (int)nodes.size();
    nodes.push_back(n);
    if (count <= 4) return idx;
    int ax = 0; float ext = -1;
    for (int k = 0; k < 3; k++) if (n.b.hi[k] - n.b.lo[k] > ext) { ext = n.b.hi[k] - n.b.lo[k]; ax = k; }
    int mid = first + count / 2;
    const std::vector<Box>& B = *boxes;
    std::nth_element(items.begin() + first, items.begin() + mid, items.begin() + first + count,
                     [&](int a, int c) { return (B[a].lo[ax] + B[a].hi[ax]) < (B[c].lo[ax] + B[c].hi[ax]); });
    int l = rec(first, mid - first), r = rec(mid, first + count - mid);
    nodes[idx].left = l; nodes[idx].right = r; nodes[idx].count = 0;
    return idx;
  }
  template <class F> void query(const Box& q, F&& fn) const {
    if (nodes.empty()) return;
    int st[128]; int sp = 0; st[sp++] = 0;
    while (sp) {
      const Node& n = nodes[st[--sp]];
      if (!overlap(q, n.b)) continue;
      if (n.count) {
        for (int i = n.first; i < n.first + n.count; i++)
          if (overlap(q, (*boxes)[items[i]])) fn(items[i]);
      }
      else { st[sp++] = n.left; st[sp++] = n.right; }
    }
  }
};

}  // namespace vf

extern "C" {

// stats: [0] faces [1] vertices [2] edges [3] boundary edges (1 face)
// [4] non-manifold edges (>2 faces) [5] orientation-inconsistent edges
// [6] non-manifold vertices [7] unreferenced vertices [8] degenerate faces
// [9] intersecting pairs [10] euler characteristic [11] bad indices
// pairs (optional): first intersecting pairs, up to pair_cap pairs.
int vf_check(const float* pts, const int* tris, int V, int F, long long stats[12], int* pairs, long long pair_cap) {
  using namespace vf;
  memset(stats, 0, 12 * sizeof(long long));
  stats[0] = F; stats[1] = V;
  for (int f = 0; f < F; f++)
    for (int k = 0; k < 3; k++)
      if (tris[3 * f + k] < 0 || tris[3 * f + k] >= V) { stats[11]++; }
  if (stats[11]) return 1;
  // edges
  std::unordered_map<unsigned long long, int> und;   // undirected -> count
  std::unordered_map<unsigned long long, int> dir;   // directed -> count
  und.reserve(3 * (size_t)F); dir.reserve(3 * (size_t)F);
  auto key = [](long long a, long long b) { return ((unsigned long long)a << 32) | (unsigned long long)b; };
  for (int f = 0; f < F; f++) {
    const int* t = &tris[3 * f];
    for (int k = 0; k < 3; k++) {
      int a = t[k], b = t[(k + 1) % 3];
      und[key(std::min(a, b), std::max(a, b))]++;
      dir[key(a, b)]++;
    }
  }
  stats[2] = (long long)und.size();
  for (auto& kv : und) {
    if (kv.second == 1) stats[3]++;
    else if (kv.second > 2) stats[4]++;
    else {
      long long a = (long long)(kv.first >> 32), b = (long long)(kv.first & 0xFFFFFFFFull);
      auto i1 = dir.find(key(a, b)), i2 = dir.find(key(b, a));
      if (i1 == dir.end() || i2 == dir.end() || i1->second != 1 || i2->second != 1) stats[5]++;
    }
  }
  // vertex links
  std::vector<int> deg(V + 1, 0);
  for (int f = 0; f < F; f++) for (int k = 0; k < 3; k++) deg[tris[3 * f + k]]++;
  std::vector<int> off(V + 1, 0);
  for (int v = 0; v < V; v++) off[v + 1] = off[v] + deg[v];
  std::vector<int> inc(off[V]);
  std::vector<int> cur(off.begin(), off.end() - 1);
  for (int f = 0; f < F; f++) for (int k = 0; k < 3; k++) inc[cur[tris[3 * f + k]]++] = f;
  for (int v = 0; v < V; v++) {
    if (deg[v] == 0) { stats[7]++; continue; }
    // link edges (b -> c) for face (v, b, c)
    std::map<int, int> nxt;
    bool bad = false;
    for (int s = off[v]; s < off[v + 1]; s++) {
      const int* t = &tris[3 * inc[s]];
      int r = t[0] == v ? 0 : (t[1] == v ? 1 : 2);
      int b = t[(r + 1) % 3], c = t[(r + 2) % 3];
      if (nxt.count(b)) bad = true;
      nxt[b] = c;
    }
    if (!bad) {
      int start = nxt.begin()->first, x = start, n = 0;
      do {
        auto it = nxt.find(x);
        if (it == nxt.end()) { bad = true; break; }
        x = it->second; n++;
      } while (x != start && n <= deg[v]);
      if (n != deg[v] || x != start) bad = true;
    }
    if (bad) stats[6]++;
  }
  stats[10] = (long long)(V - stats[7]) - stats[2] + F;
  // degenerate faces and intersections
  std::vector<P3> P(V);
  for (int v = 0; v < V; v++) P[v] = {(double)pts[3 * v], (double)pts[3 * v + 1], (double)pts[3 * v + 2]};
  std::vector<uint8_t> degen(F, 0);
  for (int f = 0; f < F; f++) {
    const int* t = &tris[3 * f];
    if (t[0] == t[1] || t[1] == t[2] || t[2] == t[0] || proj_axis(P[t[0]], P[t[1]], P[t[2]]) < 0) { degen[f] = 1; stats[8]++; }
  }
  std::vector<Box> boxes(F);
  std::vector<int> ids;
  for (int f = 0; f < F; f++) {
    if (degen[f]) continue;
    Box b;
    for (int k = 0; k < 3; k++) { b.lo[k] = INFINITY; b.hi[k] = -INFINITY; }
    for (int j = 0; j < 3; j++) {
      const float* p = &pts[3 * tris[3 * f + j]];
      for (int k = 0; k < 3; k++) { b.lo[k] = std::min(b.lo[k], p[k]); b.hi[k] = std::max(b.hi[k], p[k]); }
    }
    boxes[f] = b;
    ids.push_back(f);
  }
  BVH bvh; bvh.build(boxes, ids);
  long long nint = 0;
  std::vector<std::pair<int, int>> found;
#pragma omp parallel
  {
    std::vector<std::pair<int, int>> local;
#pragma omp for schedule(dynamic, 256) reduction(+ : nint)
    for (size_t ii = 0; ii < ids.size(); ii++) {
      int f = ids[ii];
      const int* t = &tris[3 * f];
      bvh.query(boxes[f], [&](int g) {
        if (g <= f) return;
        const int* u = &tris[3 * g];
        int shared = 0, ti[3] = {-1, -1, -1}, ui[3] = {-1, -1, -1};
        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) if (t[a] == u[b]) { ti[shared] = a; ui[shared] = b; shared++; }
        bool hit;
        if (shared >= 3) hit = true;
        else if (shared == 2) {
          int ta = 3 - ti[0] - ti[1], ua = 3 - ui[0] - ui[1];
          hit = tri_tri_shared_edge(P[t[ti[0]]], P[t[ti[1]]], P[t[ta]], P[u[ua]]);
        } else if (shared == 1) {
          int p = t[ti[0]];
          int a = t[(ti[0] + 1) % 3], b = t[(ti[0] + 2) % 3];
          int c = u[(ui[0] + 1) % 3], d = u[(ui[0] + 2) % 3];
          hit = tri_tri_shared_vertex(P[p], P[a], P[b], P[c], P[d]);
        } else {
          P3 A[3] = {P[t[0]], P[t[1]], P[t[2]]}, B[3] = {P[u[0]], P[u[1]], P[u[2]]};
          hit = tri_tri_general(A, B);
        }
        if (hit) { nint++; if ((long long)local.size() < pair_cap) local.push_back({f, g}); }
      });
    }
#pragma omp critical
    found.insert(found.end(), local.begin(), local.end());
  }
  stats[9] = nint;
  std::sort(found.begin(), found.end());
  for (long long i = 0; i < (long long)found.size() && i < pair_cap; i++) { pairs[2 * i] = found[i].first; pairs[2 * i + 1] = found[i].second; }
  return 0;
}

int vf_orient3d(const double a[3], const double b[3], const double c[3], const double d[3]) {
  vf::P3 A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]}, D{d[0], d[1], d[2]};
  return vf::orient3d(A, B, C, D);
}
int vf_orient3d_exact(const double a[3], const double b[3], const double c[3], const double d[3]) {
  vf::P3 A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]}, D{d[0], d[1], d[2]};
  return vf::orient3d_exact(A, B, C, D);
}
// exact closed triangle-triangle test honouring shared vertex indices
int vf_tri_pair(const float a[9], const float b[9], const int ia[3], const int ib[3]) {
  float pts[18];
  memcpy(pts, a, 36); memcpy(pts + 9, b, 36);
  // build a 2-face mesh whose vertex ids follow ia/ib
  int maxid = 0;
  for (int k = 0; k < 3; k++) maxid = std::max(maxid, std::max(ia[k], ib[k]));
  std::vector<float> P((size_t)3 * (maxid + 1), 0.0f);
  for (int k = 0; k < 3; k++) { memcpy(&P[3 * ia[k]], a + 3 * k, 12); memcpy(&P[3 * ib[k]], b + 3 * k, 12); }
  int T[6] = {ia[0], ia[1], ia[2], ib[0], ib[1], ib[2]};
  long long st[12]; int pr[2];
  vf_check(P.data(), T, maxid + 1, 2, st, pr, 1);
  return st[9] > 0 ? 1 : 0;
}

}  // extern "C"
