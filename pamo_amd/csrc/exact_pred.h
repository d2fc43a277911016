// Exact orientation predicates and the exact closed triangle-triangle test
// on float coordinates, shared by the host verifier (verify.cpp), the CPU
// oracle's strict guard and the device guard's strict mode (hipcc).
//
// Predicates: orient3d / orient2d evaluated in double with Shewchuk's static
// error bound and, when that filter cannot decide, an exact evaluation in
// expansion arithmetic (J. R. Shewchuk, "Adaptive Precision Floating-Point
// Arithmetic and Fast Robust Geometric Predicates", 1997: Two-Sum,
// Two-Diff, Dekker split / Two-Product, Scale-Expansion, and Grow-Expansion
// summation with zero elimination).  Expansions live in fixed arrays sized
// for the worst case of each step (differences 2 terms, products of two
// differences 8, 2x2 minors 16, their products with a difference 64, the
// determinant 192), so the code has no allocation and runs on the device.
// Compile with -ffp-contract=off.
//
// Intersection semantics (closed triangles, as CGAL's do_intersect):
//   * no shared vertex index: any common point is an intersection;
//   * one shared vertex p: an intersection iff the triangles share a point
//     other than p;
//   * one shared edge: an intersection iff the triangles are coplanar and lie
//     on the same side of the edge (they fold over each other);
//   * three shared indices (duplicate face): an intersection.
// Degenerate faces (a repeated index, or collinear corners) are not tested.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define XP_FN __host__ __device__ inline
#define XP_NOINLINE __host__ __device__ __attribute__((noinline))
#else
#define XP_FN inline
#define XP_NOINLINE __attribute__((noinline))
#endif

namespace xp {

// ------------------------------------------------ expansion arithmetic
XP_FN void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  double bv = x - a, av = x - bv, br = b - bv, ar = a - av;
  y = ar + br;
}
XP_FN void two_diff(double a, double b, double& x, double& y) {
  x = a - b;
  double bv = a - x, av = x + bv, br = bv - b, ar = a - av;
  y = ar + br;
}
XP_FN void split(double a, double& hi, double& lo) {
  const double splitter = 134217729.0;   // 2^27 + 1
  double c = splitter * a, abig = c - a;
  hi = c - abig;
  lo = a - hi;
}
XP_FN void two_product(double a, double b, double& x, double& y) {
  x = a * b;
  double ahi, alo, bhi, blo;
  split(a, ahi, alo);
  split(b, bhi, blo);
  double err1 = x - ahi * bhi, err2 = err1 - alo * bhi, err3 = err2 - ahi * blo;
  y = alo * blo - err3;
}
// h = e + f (both nonoverlapping, increasing magnitude), by one
// Grow-Expansion per component of f; h holds at least ne + nf terms.
// Returns the length (at least 1).
XP_FN int exp_sum(const double* e, int ne, const double* f, int nf, double* h) {
  int n = 0;
  for (int i = 0; i < ne; i++) h[n++] = e[i];
  if (n == 0) h[n++] = 0.0;
  for (int j = 0; j < nf; j++) {
    double q = f[j];
    int m = 0;
    for (int i = 0; i < n; i++) {
      double x, y;
      two_sum(q, h[i], x, y);
      if (y != 0.0) h[m++] = y;
      q = x;
    }
    if (q != 0.0 || m == 0) h[m++] = q;
    n = m;
  }
  return n;
}
// h = e * b; h holds at least 2 ne terms.  Returns the length (at least 1).
XP_FN int exp_scale(const double* e, int ne, double b, double* h) {
  int n = 0;
  double q, y;
  two_product(e[0], b, q, y);
  if (y != 0.0) h[n++] = y;
  for (int i = 1; i < ne; i++) {
    double p1, p0, s, t;
    two_product(e[i], b, p1, p0);
    two_sum(q, p0, s, t);
    if (t != 0.0) h[n++] = t;
    two_sum(p1, s, q, t);
    if (t != 0.0) h[n++] = t;
  }
  if (q != 0.0 || n == 0) h[n++] = q;
  return n;
}
XP_FN int exp_sign(const double* e, int n) {
  for (int i = n - 1; i >= 0; i--)
    if (e[i] != 0.0) return e[i] > 0 ? 1 : -1;
  return 0;
}
XP_FN int diff2(double a, double b, double* e) {   // a - b, 1 or 2 terms
  double x, y;
  two_diff(a, b, x, y);
  if (y != 0.0) { e[0] = y; e[1] = x; return 2; }
  e[0] = x;
  return 1;
}
// h = a * b for a, b of at most 2 terms: at most 8 terms
XP_FN int mul2x2(const double* a, int na, const double* b, int nb, double* h) {
  double s[4], t[4];
  int ns = exp_scale(a, na, b[0], s);
  if (nb == 1) {
    for (int i = 0; i < ns; i++) h[i] = s[i];
    return ns;
  }
  int nt = exp_scale(a, na, b[1], t);
  return exp_sum(s, ns, t, nt, h);
}
// h = a * m for a of at most 2 terms and m of at most 16: at most 64 terms
XP_FN int mul2x16(const double* a, int na, const double* m, int nm, double* h) {
  double s[32], t[32];
  int ns = exp_scale(m, nm, a[0], s);
  if (na == 1) {
    for (int i = 0; i < ns; i++) h[i] = s[i];
    return ns;
  }
  int nt = exp_scale(m, nm, a[1], t);
  return exp_sum(s, ns, t, nt, h);
}
// h = p*q - r*s for differences of at most 2 terms: at most 16 terms
XP_FN int minor2(const double* p, int np, const double* q, int nq, const double* r, int nr, const double* s, int ns,
                 double* h) {
  double a[8], b[8];
  int na = mul2x2(p, np, q, nq, a);
  int nb = mul2x2(r, nr, s, ns, b);
  for (int i = 0; i < nb; i++) b[i] = -b[i];
  return exp_sum(a, na, b, nb, h);
}

struct P3 { double x, y, z; };
struct P2 { double x, y; };

XP_NOINLINE int orient3d_exact(const P3& a, const P3& b, const P3& c, const P3& d) {
  double adx[2], ady[2], adz[2], bdx[2], bdy[2], bdz[2], cdx[2], cdy[2], cdz[2];
  int nadx = diff2(a.x, d.x, adx), nady = diff2(a.y, d.y, ady), nadz = diff2(a.z, d.z, adz);
  int nbdx = diff2(b.x, d.x, bdx), nbdy = diff2(b.y, d.y, bdy), nbdz = diff2(b.z, d.z, bdz);
  int ncdx = diff2(c.x, d.x, cdx), ncdy = diff2(c.y, d.y, cdy), ncdz = diff2(c.z, d.z, cdz);
  double t[16], u1[64], u2[64], u3[64], s12[128], det[192];
  int nt = minor2(bdx, nbdx, cdy, ncdy, bdy, nbdy, cdx, ncdx, t);
  int n1 = mul2x16(adz, nadz, t, nt, u1);
  nt = minor2(cdx, ncdx, ady, nady, cdy, ncdy, adx, nadx, t);
  int n2 = mul2x16(bdz, nbdz, t, nt, u2);
  nt = minor2(adx, nadx, bdy, nbdy, ady, nady, bdx, nbdx, t);
  int n3 = mul2x16(cdz, ncdz, t, nt, u3);
  int n12 = exp_sum(u1, n1, u2, n2, s12);
  int nd = exp_sum(s12, n12, u3, n3, det);
  return exp_sign(det, nd);
}
// sign of det[a-d; b-d; c-d]
XP_FN int orient3d(const P3& a, const P3& b, const P3& c, const P3& d) {
  const double eps = 1.1102230246251565e-16;   // 2^-53
  const double o3d_bound = (7.0 + 56.0 * eps) * eps;
  double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
  double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
  double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
  double perm = (std::fabs(bdxcdy) + std::fabs(cdxbdy)) * std::fabs(adz) +
                (std::fabs(cdxady) + std::fabs(adxcdy)) * std::fabs(bdz) +
                (std::fabs(adxbdy) + std::fabs(bdxady)) * std::fabs(cdz);
  double bound = o3d_bound * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return orient3d_exact(a, b, c, d);
}

XP_NOINLINE int orient2d_exact(const P2& a, const P2& b, const P2& c) {
  double acx[2], acy[2], bcx[2], bcy[2], h[16];
  int n1 = diff2(a.x, c.x, acx), n2 = diff2(a.y, c.y, acy), n3 = diff2(b.x, c.x, bcx), n4 = diff2(b.y, c.y, bcy);
  int n = minor2(acx, n1, bcy, n4, acy, n2, bcx, n3, h);
  return exp_sign(h, n);
}
XP_FN int orient2d(const P2& a, const P2& b, const P2& c) {
  const double eps = 1.1102230246251565e-16;
  const double ccw_bound = (3.0 + 16.0 * eps) * eps;
  double dl = (a.x - c.x) * (b.y - c.y), dr = (a.y - c.y) * (b.x - c.x);
  double det = dl - dr;
  double bound = ccw_bound * (std::fabs(dl) + std::fabs(dr));
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return orient2d_exact(a, b, c);
}

XP_FN P2 proj(const P3& p, int drop) {
  if (drop == 0) return {p.y, p.z};
  if (drop == 1) return {p.x, p.z};
  return {p.x, p.y};
}
// axis to drop so that the triangle's projection is non-degenerate; -1 if
// the triangle is degenerate (collinear)
XP_FN int proj_axis(const P3& a, const P3& b, const P3& c) {
  for (int drop = 2; drop >= 0; drop--)
    if (orient2d(proj(a, drop), proj(b, drop), proj(c, drop)) != 0) return drop;
  return -1;
}

// --------------------------------------------------------- 2-D tests
XP_FN double dmin(double a, double b) { return a < b ? a : b; }
XP_FN double dmax(double a, double b) { return a > b ? a : b; }
XP_FN bool on_seg_2d(const P2& p, const P2& q, const P2& r) {   // r collinear with pq
  return dmin(p.x, q.x) <= r.x && r.x <= dmax(p.x, q.x) && dmin(p.y, q.y) <= r.y && r.y <= dmax(p.y, q.y);
}
XP_FN bool seg_seg_2d(const P2& p1, const P2& p2, const P2& q1, const P2& q2) {
  int d1 = orient2d(q1, q2, p1), d2 = orient2d(q1, q2, p2), d3 = orient2d(p1, p2, q1), d4 = orient2d(p1, p2, q2);
  if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return true;
  if (d1 == 0 && on_seg_2d(q1, q2, p1)) return true;
  if (d2 == 0 && on_seg_2d(q1, q2, p2)) return true;
  if (d3 == 0 && on_seg_2d(p1, p2, q1)) return true;
  if (d4 == 0 && on_seg_2d(p1, p2, q2)) return true;
  return false;
}
XP_FN bool pt_in_tri_2d(const P2& p, const P2& a, const P2& b, const P2& c) {
  int s = orient2d(a, b, c);
  int o1 = orient2d(a, b, p) * s, o2 = orient2d(b, c, p) * s, o3 = orient2d(c, a, p) * s;
  return o1 >= 0 && o2 >= 0 && o3 >= 0;
}
XP_FN bool seg_tri_2d(const P2& p, const P2& q, const P2& a, const P2& b, const P2& c) {
  if (pt_in_tri_2d(p, a, b, c) || pt_in_tri_2d(q, a, b, c)) return true;
  return seg_seg_2d(p, q, a, b) || seg_seg_2d(p, q, b, c) || seg_seg_2d(p, q, c, a);
}
XP_FN bool tri_tri_2d(const P2 t[3], const P2 u[3]) {
  for (int i = 0; i < 3; i++)
    if (pt_in_tri_2d(t[i], u[0], u[1], u[2]) || pt_in_tri_2d(u[i], t[0], t[1], t[2])) return true;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (seg_seg_2d(t[i], t[(i + 1) % 3], u[j], u[(j + 1) % 3])) return true;
  return false;
}

// --------------------------------------------------------- 3-D tests
// closed segment pq against closed triangle abc
XP_FN bool seg_tri(const P3& p, const P3& q, const P3& a, const P3& b, const P3& c) {
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

XP_FN bool tri_tri_general(const P3 t[3], const P3 u[3]) {
  int o[3], w[3];
  for (int i = 0; i < 3; i++) o[i] = orient3d(t[0], t[1], t[2], u[i]);
  if ((o[0] > 0 && o[1] > 0 && o[2] > 0) || (o[0] < 0 && o[1] < 0 && o[2] < 0)) return false;
  for (int i = 0; i < 3; i++) w[i] = orient3d(u[0], u[1], u[2], t[i]);
  if ((w[0] > 0 && w[1] > 0 && w[2] > 0) || (w[0] < 0 && w[1] < 0 && w[2] < 0)) return false;
  if (o[0] == 0 && o[1] == 0 && o[2] == 0) {
    int ax = proj_axis(t[0], t[1], t[2]);
    if (ax < 0) return false;
    P2 a[3] = {proj(t[0], ax), proj(t[1], ax), proj(t[2], ax)};
    P2 b[3] = {proj(u[0], ax), proj(u[1], ax), proj(u[2], ax)};
    return tri_tri_2d(a, b);
  }
  for (int i = 0; i < 3; i++)
    if (seg_tri(t[i], t[(i + 1) % 3], u[0], u[1], u[2])) return true;
  for (int i = 0; i < 3; i++)
    if (seg_tri(u[i], u[(i + 1) % 3], t[0], t[1], t[2])) return true;
  return false;
}

// segment p->a (p a vertex of triangle p,c,d) meets the triangle in a point
// other than p
XP_FN bool ray_from_shared(const P3& p, const P3& a, const P3& c, const P3& d) {
  if (orient3d(p, c, d, a) != 0) return false;
  int ax = proj_axis(p, c, d);
  if (ax < 0) return false;
  P2 P = proj(p, ax), A = proj(a, ax), C = proj(c, ax), D = proj(d, ax);
  int s = orient2d(P, C, D);
  return s * orient2d(P, C, A) >= 0 && s * orient2d(P, A, D) >= 0;
}

// triangles (p,a,b) and (p,c,d) sharing exactly vertex p
XP_FN bool tri_tri_shared_vertex(const P3& p, const P3& a, const P3& b, const P3& c, const P3& d) {
  if (seg_tri(a, b, p, c, d)) return true;
  if (seg_tri(c, d, p, a, b)) return true;
  if (ray_from_shared(p, a, c, d) || ray_from_shared(p, b, c, d)) return true;
  if (ray_from_shared(p, c, a, b) || ray_from_shared(p, d, a, b)) return true;
  return false;
}

// triangles (p,q,a) and (p,q,b) sharing edge pq
XP_FN bool tri_tri_shared_edge(const P3& p, const P3& q, const P3& a, const P3& b) {
  if (orient3d(p, q, a, b) != 0) return false;
  int ax = proj_axis(p, q, a);
  if (ax < 0) return false;
  int sa = orient2d(proj(p, ax), proj(q, ax), proj(a, ax));
  int sb = orient2d(proj(p, ax), proj(q, ax), proj(b, ax));
  return sa * sb >= 0;
}

// a face is degenerate when an index repeats or its corners are collinear
XP_FN bool degenerate(const P3 t[3], const int it[3]) {
  return it[0] == it[1] || it[1] == it[2] || it[2] == it[0] || proj_axis(t[0], t[1], t[2]) < 0;
}

// Exact closed intersection of faces t (indices it) and u (indices iu) with
// the semantics above; degenerate faces never intersect here.
XP_FN bool tri_pair(const P3 t[3], const int it[3], const P3 u[3], const int iu[3]) {
  if (degenerate(t, it) || degenerate(u, iu)) return false;
  int shared = 0, ti[3] = {-1, -1, -1}, ui[3] = {-1, -1, -1};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (it[a] == iu[b] && shared < 3) { ti[shared] = a; ui[shared] = b; shared++; }
  if (shared >= 3) return true;
  if (shared == 2) {
    int ta = 3 - ti[0] - ti[1], ua = 3 - ui[0] - ui[1];
    return tri_tri_shared_edge(t[ti[0]], t[ti[1]], t[ta], u[ua]);
  }
  if (shared == 1) {
    const P3& p = t[ti[0]];
    return tri_tri_shared_vertex(p, t[(ti[0] + 1) % 3], t[(ti[0] + 2) % 3], u[(ui[0] + 1) % 3], u[(ui[0] + 2) % 3]);
  }
  return tri_tri_general(t, u);
}

}  // namespace xp
