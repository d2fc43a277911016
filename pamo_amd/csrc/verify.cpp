// EXACT MESH VERIFIER (libpamo_verify.so; CLI: python -m pamo_amd.verify).
//
// Checks a triangle mesh for (1) manifoldness and watertightness,
// (2) degenerate faces and (3) self-intersections, with exact predicates.
// It replaces the reference's offline CGAL EPICK checker
// (simp_cuda/safe_project/scripts/cgal_intersect/main.cpp:12,66), which is
// not buildable here (no CGAL/GMP).
//
// Predicates and the pair test come from exact_pred.h (orient3d / orient2d
// with Shewchuk's filters and exact expansion arithmetic; the closed
// triangle-triangle semantics of CGAL's do_intersect), which the device
// guard's strict mode and the oracle share.  Degenerate (collinear or
// repeated-index) faces are counted separately and skipped by the
// intersection test.
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

#include "exact_pred.h"

namespace vf {

using xp::P2;
using xp::P3;
using xp::proj_axis;
using xp::tri_pair;

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
        const int it[3] = {t[0], t[1], t[2]}, iu[3] = {u[0], u[1], u[2]};
        const P3 A[3] = {P[t[0]], P[t[1]], P[t[2]]}, B[3] = {P[u[0]], P[u[1]], P[u[2]]};
        bool hit = tri_pair(A, it, B, iu);
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
  return xp::orient3d(A, B, C, D);
}
int vf_orient3d_exact(const double a[3], const double b[3], const double c[3], const double d[3]) {
  vf::P3 A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]}, D{d[0], d[1], d[2]};
  return xp::orient3d_exact(A, B, C, D);
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
