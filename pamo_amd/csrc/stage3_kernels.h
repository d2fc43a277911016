// Stage 3 (safe projection) on the device, in float64: the Newton / CG /
// CCD loop of simp_cuda/safe_project/src/pamo_safe_project/system.py:17-470
// with the four energies of the default configuration (config.py:26-35):
//   target distance   energy.py:193-211, distance_energy.py:6-47,
//                     geometry_kernels.py:91-104
//   membrane          energy.py:337-432, elastic_energy.py:1-220
//   hinge bending     energy.py:589-700, hinge_energy.py:1-274,
//                     utils_kernels.py:20-41
//   contact barrier   energy.py:705-946, collision_energy.py:1-409,
//                     contact_detection.py:120-216, ccd_kernels.py:8-110
// and the distance classification of
// kernels/distance_kernels/distance_kernels.py:1-545.
//
// The reference computes in float32 with float atomics; here every
// quantity is float64 and every sum has a fixed order:
//  * per-vertex sums (forces, diagonal, H dx) are gathers over a vertex ->
//    element-slot list sorted by element (its faces, hinges and contacts,
//    each in ascending order): 32 lanes add every 32nd entry in order, a
//    butterfly adds the lanes, then the vertex's own term is added; Voronoi
//    areas are serial sums over the vertex's faces in ascending order;
//  * global sums (energies, CG dot products) are taken by one workgroup:
//    lane t sums elements t, t + 1024, ... in order, then a fixed tree.
// Contacts are listed point-triangle pairs first, by (vertex, triangle
// corners), then edge-edge pairs by (edge, edge), edges numbered in
// ascending (min, max) vertex order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace s3 {

// ------------------------------------------------------------- vectors
struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 mk3(double x, double y, double z) {
  D3 r;
  r.x = x; r.y = y; r.z = z;
  return r;
}
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator-(D3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ D3 operator/(D3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double len2(D3 a) { return dot(a, a); }
__device__ __forceinline__ double len(D3 a) { return sqrt(dot(a, a)); }
// the zero vector normalises to zero (wp.normalize)
__device__ __forceinline__ D3 nrm(D3 a) {
  const double l = len(a);
  return l > 0.0 ? a / l : mk3(0.0, 0.0, 0.0);
}
__device__ __forceinline__ D3 cmul(D3 a, D3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ D3 ld3(const double* __restrict__ p, long long i) {
  return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ void st3(double* __restrict__ p, long long i, D3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}

// ---------------------------------------------------------- distances
// contact types (defs.py:1-31)
enum { PT = 0, PP01 = 1, PP02 = 2, PE012 = 3, PP03 = 4, PE031 = 5, PE023 = 6 };
enum { EPP02 = 0, EPP03 = 1, EPP12 = 2, EPP13 = 3, EPE023 = 4, EPE123 = 5, EPE201 = 6, EPE301 = 7, EEE = 8 };
enum { BT_PT = 3, BT_EE = 4 };

// distance_kernels.py:8-27
__device__ __forceinline__ double d_pt_plane(D3 x0, D3 x1, D3 x2, D3 x3) {
  return fabs(dot(x0 - x1, nrm(cross(x2 - x1, x3 - x1))));
}
__device__ __forceinline__ double d_ee_line(D3 x0, D3 x1, D3 x2, D3 x3) {
  return fabs(dot(x0 - x2, nrm(cross(x1 - x0, x3 - x2))));
}
__device__ __forceinline__ double d_pe(D3 x0, D3 x1, D3 x2) { return len(cross(x1 - x0, x2 - x0)) / len(x1 - x2); }
__device__ __forceinline__ double d_pp(D3 x0, D3 x1) { return len(x1 - x0); }

// Point x0 against triangle (x1, x2, x3): which feature is closest
// (distance_kernels.py:79-205).  Nearly degenerate triangles use the nearest
// corner; otherwise the barycentric signs of the projection pick the face,
// an edge or a corner.
__device__ int pt_classify(D3 x0, D3 x1, D3 x2, D3 x3) {
  D3 x10 = x0 - x1;
  const D3 x12 = x2 - x1, x13 = x3 - x1;
  const double area_sq = len2(cross(x12, x13));
  if (area_sq <= 1e-6 * (len2(x12) * len2(x13)) || area_sq <= 1e-16) {
    const double d1 = len(x10), d2 = len(x0 - x2), d3 = len(x0 - x3);
    if (d1 < d2) return d1 < d3 ? PP01 : PP03;
    return d2 < d3 ? PP02 : PP03;
  }
  const D3 n = nrm(cross(x12, x13));
  const D3 xp = x0 - dot(n, x10) * n;
  x10 = xp - x1;
  const D3 x20 = xp - x2, x30 = xp - x3;
  const double a00 = dot(x12, x12), a01 = dot(x12, x13), a11 = dot(x13, x13);
  const double b0 = dot(x10, x12), b1 = dot(x10, x13);
  const double det = a00 * a11 - a01 * a01;
  const double beta = b0 * a11 - b1 * a01, gamma = a00 * b1 - b0 * a01, alpha = det - beta - gamma;
  const int nneg = (alpha < 0.0) + (beta < 0.0) + (gamma < 0.0);
  if (nneg == 0) return PT;
  if (nneg == 1) {
    if (alpha < 0.0) {   // beyond edge x2x3: the unprojected point against it
      const D3 x23 = x3 - x2;
      const double r = dot(x23, x0 - x2);
      if (r < 0.0) return PP02;
      if (r > dot(x23, x23)) return PP03;
      return PE023;
    }
    if (beta < 0.0) {   // beyond edge x3x1
      const D3 x31 = x1 - x3;
      const double r = dot(x31, x0 - x3);
      if (r < 0.0) return PP03;
      if (r > dot(x31, x31)) return PP01;
      return PE031;
    }
    const double r = dot(x12, x10);   // beyond edge x1x2, projected point
    if (r < 0.0) return PP01;
    if (r > dot(x12, x12)) return PP02;
    return PE012;
  }
  if (alpha >= 0.0) {   // beyond corner x1
    const double d12 = dot(x12, x10), d13 = dot(x13, x10);
    if (d12 > 0.0 && d13 <= 0.0) return dot(x12, x20) > 0.0 ? PP02 : PE012;
    if (d13 > 0.0 && d12 <= 0.0) return dot(x13, x30) > 0.0 ? PP03 : PE031;
    return PP01;
  }
  if (beta >= 0.0) {   // beyond corner x2
    const D3 x23 = x3 - x2, x21 = x1 - x2;
    const double d23 = dot(x23, x20), d21 = dot(x21, x20);
    if (d23 > 0.0 && d21 <= 0.0) return dot(x23, x30) > 0.0 ? PP03 : PE023;
    if (d21 > 0.0 && d23 <= 0.0) return dot(x21, x10) > 0.0 ? PP01 : PE012;
    return PP02;
  }
  const D3 x31 = x1 - x3, x32 = x2 - x3;   // beyond corner x3
  const double d31 = dot(x31, x30), d32 = dot(x32, x30);
  if (d31 > 0.0 && d32 <= 0.0) return dot(x31, x10) > 0.0 ? PP01 : PE031;
  if (d32 > 0.0 && d31 <= 0.0) return dot(x32, x20) > 0.0 ? PP02 : PE023;
  return PP03;
}

// distance_kernels.py:228-245
__device__ double pt_distance(D3 x0, D3 x1, D3 x2, D3 x3, int t) {
  switch (t) {
    case PT: return d_pt_plane(x0, x1, x2, x3);
    case PP01: return d_pp(x0, x1);
    case PP02: return d_pp(x0, x2);
    case PE012: return d_pe(x0, x1, x2);
    case PP03: return d_pp(x0, x3);
    case PE031: return d_pe(x0, x3, x1);
    case PE023: return d_pe(x0, x2, x3);
  }
  return 1e9;
}

// Gradients of those distances.  The distance to an affine set (point,
// line, plane) is |x0 - y*| at the foot y*; moving a point that defines the
// set moves the foot with weight equal to that point's affine coordinate
// of y*, so the gradient is the unit direction u = (x0 - y*)/|x0 - y*| on
// x0 and -w_k u on each defining point k.
__device__ __forceinline__ void g_pp(D3 a, D3 b, int ia, int ib, D3 g[4]) {
  g[ia] = nrm(a - b);
  g[ib] = nrm(b - a);
}
__device__ __forceinline__ void g_pe(D3 p, D3 a, D3 b, int ip, int ia, int ib, D3 g[4]) {
  const D3 ab = b - a;
  const double t = dot(p - a, ab) / dot(ab, ab);
  const D3 u = nrm(p - (a + t * ab));
  g[ip] = u;
  g[ia] = -(1.0 - t) * u;
  g[ib] = -t * u;
}
__device__ __forceinline__ void g_pt_plane(D3 x0, D3 x1, D3 x2, D3 x3, D3 g[4]) {
  const D3 x12 = x2 - x1, x13 = x3 - x1, x10 = x0 - x1;
  const D3 n = nrm(cross(x12, x13));
  const double s = dot(x10, n);
  const D3 u = s > 0.0 ? n : (s < 0.0 ? -n : mk3(0.0, 0.0, 0.0));
  const double a00 = dot(x12, x12), a01 = dot(x12, x13), a11 = dot(x13, x13);
  const double b0 = dot(x10, x12), b1 = dot(x10, x13);
  const double det = a00 * a11 - a01 * a01;
  const double beta = (b0 * a11 - b1 * a01) / det, gamma = (a00 * b1 - b0 * a01) / det;
  const double alpha = 1.0 - beta - gamma;
  g[0] = u;
  g[1] = -alpha * u;
  g[2] = -beta * u;
  g[3] = -gamma * u;
}
__device__ __forceinline__ void g_ee_line(D3 x0, D3 x1, D3 x2, D3 x3, D3 g[4]) {
  const D3 u = x1 - x0, v = x3 - x2, w = x0 - x2;
  const D3 n = nrm(cross(u, v));
  const double s = dot(w, n);
  const D3 un = s > 0.0 ? n : (s < 0.0 ? -n : mk3(0.0, 0.0, 0.0));
  const double A = dot(u, u), B = dot(u, v), C = dot(v, v), D = dot(u, w), E = dot(v, w);
  const double den = A * C - B * B;
  const double a = (B * E - C * D) / den, b = (A * E - B * D) / den;
  g[0] = (1.0 - a) * un;
  g[1] = a * un;
  g[2] = -(1.0 - b) * un;
  g[3] = -b * un;
}
__device__ void pt_grad(D3 x0, D3 x1, D3 x2, D3 x3, int t, D3 g[4]) {
  for (int k = 0; k < 4; k++) g[k] = mk3(0.0, 0.0, 0.0);
  switch (t) {
    case PT: g_pt_plane(x0, x1, x2, x3, g); break;
    case PP01: g_pp(x0, x1, 0, 1, g); break;
    case PP02: g_pp(x0, x2, 0, 2, g); break;
    case PE012: g_pe(x0, x1, x2, 0, 1, 2, g); break;
    case PP03: g_pp(x0, x3, 0, 3, g); break;
    case PE031: g_pe(x0, x3, x1, 0, 3, 1, g); break;
    case PE023: g_pe(x0, x2, x3, 0, 2, 3, g); break;
  }
}

// Edge (x0, x1) against edge (x2, x3) (distance_kernels.py:309-430): the
// interior pair when both line parameters fall inside and the edges are
// not nearly parallel, else the nearest of the four endpoint-to-edge
// features (the first of equal ones).
__device__ int ee_classify(D3 x0, D3 x1, D3 x2, D3 x3, double sin_thres) {
  const D3 e0 = x0 - x1, e1 = x2 - x3;
  const double e0s = dot(e0, e0), e1s = dot(e1, e1);
  const double a00 = e0s, a01 = -dot(e0, e1), a10 = dot(e0, e1), a11 = -e1s;
  const double b0 = dot(x3 - x1, e0), b1 = dot(x3 - x1, e1);
  const double mu0 = -(b0 * a11 - b1 * a01), mu1 = -(a00 * b1 - a10 * b0), detA = -(a00 * a11 - a01 * a10);
  const double esin = len(cross(nrm(e0), nrm(e1)));
  if (mu0 >= 0 && mu0 <= detA && mu1 >= 0 && mu1 <= detA && esin >= sin_thres && e0s > 1e-15 && e1s > 1e-15)
    return EEE;
  double best = 1e9;
  int type = -1;
#define S3_TAKE(d_, t_)      \
  {                          \
    const double dd_ = (d_); \
    if (dd_ < best) {        \
      best = dd_;            \
      type = (t_);           \
    }                        \
  }
  const double t0 = dot(x0 - x3, e1);
  if (t0 >= e1s) S3_TAKE(d_pp(x0, x2), EPP02) else if (t0 <= 0) S3_TAKE(d_pp(x0, x3), EPP03) else S3_TAKE(d_pe(x0, x2, x3), EPE023)
  const double t1 = dot(x1 - x3, e1);
  if (t1 >= e1s) S3_TAKE(d_pp(x1, x2), EPP12) else if (t1 <= 0) S3_TAKE(d_pp(x1, x3), EPP13) else S3_TAKE(d_pe(x1, x2, x3), EPE123)
  const double t2 = dot(x2 - x1, e0);
  if (t2 >= e0s) S3_TAKE(d_pp(x2, x0), EPP02) else if (t2 <= 0) S3_TAKE(d_pp(x2, x1), EPP12) else S3_TAKE(d_pe(x2, x0, x1), EPE201)
  const double t3 = dot(x3 - x1, e0);
  if (t3 >= e0s) S3_TAKE(d_pp(x3, x0), EPP03) else if (t3 <= 0) S3_TAKE(d_pp(x3, x1), EPP13) else S3_TAKE(d_pe(x3, x0, x1), EPE301)
#undef S3_TAKE
  return type;
}

// distance_kernels.py:433-452
__device__ double ee_distance(D3 x0, D3 x1, D3 x2, D3 x3, int t) {
  switch (t) {
    case EPP02: return d_pp(x0, x2);
    case EPP03: return d_pp(x0, x3);
    case EPP12: return d_pp(x1, x2);
    case EPP13: return d_pp(x1, x3);
    case EPE023: return d_pe(x0, x2, x3);
    case EPE123: return d_pe(x1, x2, x3);
    case EPE201: return d_pe(x2, x0, x1);
    case EPE301: return d_pe(x3, x0, x1);
    case EEE: return d_ee_line(x0, x1, x2, x3);
  }
  return 1e9;
}
__device__ void ee_grad(D3 x0, D3 x1, D3 x2, D3 x3, int t, D3 g[4]) {
  for (int k = 0; k < 4; k++) g[k] = mk3(0.0, 0.0, 0.0);
  switch (t) {
    case EPP02: g_pp(x0, x2, 0, 2, g); break;
    case EPP03: g_pp(x0, x3, 0, 3, g); break;
    case EPP12: g_pp(x1, x2, 1, 2, g); break;
    case EPP13: g_pp(x1, x3, 1, 3, g); break;
    case EPE023: g_pe(x0, x2, x3, 0, 2, 3, g); break;
    case EPE123: g_pe(x1, x2, x3, 1, 2, 3, g); break;
    case EPE201: g_pe(x2, x0, x1, 2, 0, 1, g); break;
    case EPE301: g_pe(x3, x0, x1, 3, 0, 1, g); break;
    case EEE: g_ee_line(x0, x1, x2, x3, g); break;
  }
}

// Closest point of triangle (a, b, c) to p by Voronoi regions (Ericson,
// Real-Time Collision Detection, 5.1.5): the query of
// wp.mesh_query_point_no_sign + mesh_eval_position (geometry_kernels.py:91-104).
__device__ D3 closest_on_tri(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const D3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + (d1 / (d1 - d3)) * ab;
  const D3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + (d2 / (d2 - d6)) * ac;
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
  const double den = 1.0 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  return a + v * ab + w * ac;
}

// IPC log barrier and its derivatives (collision_energy.py:8-27)
__device__ __forceinline__ double bar(double d, double dh) {
  return d >= dh ? 0.0 : -(d - dh) * (d - dh) * log(d / dh);
}
__device__ __forceinline__ double bar1(double d, double dh) {
  const double t = d - dh;
  return d >= dh ? 0.0 : -2.0 * t * log(d / dh) - t * t / d;
}
__device__ __forceinline__ double bar2(double d, double dh) {
  return d >= dh ? 0.0 : -2.0 * log(d / dh) - 2.0 * (d - dh) / d - 1.0 + (dh * dh) / (d * d);
}

// ----------------------------------------------------------- membrane
// Green strain St. Venant-Kirchhoff triangles in their own 2D frame
// (elastic_energy.py:1-220).  2x2 row-major (a b; c d).
struct M2 {
  double a, b, c, d;
};
__device__ __forceinline__ M2 m2(double a, double b, double c, double d) {
  M2 r;
  r.a = a; r.b = b; r.c = c; r.d = d;
  return r;
}
__device__ __forceinline__ M2 mul(M2 x, M2 y) {
  return m2(x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d);
}
__device__ __forceinline__ M2 tr(M2 x) { return m2(x.a, x.c, x.b, x.d); }
__device__ __forceinline__ M2 add(M2 x, M2 y) { return m2(x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d); }
__device__ __forceinline__ M2 scl(double s, M2 x) { return m2(s * x.a, s * x.b, s * x.c, s * x.d); }
__device__ __forceinline__ double trace(M2 x) { return x.a + x.d; }
// wp.inverse: the zero matrix for a zero determinant
__device__ __forceinline__ M2 inv(M2 x) {
  const double det = x.a * x.d - x.b * x.c;
  if (det == 0.0) return m2(0.0, 0.0, 0.0, 0.0);
  return m2(x.d / det, -x.b / det, -x.c / det, x.a / det);
}
__device__ __forceinline__ void frame(D3 x0, D3 x1, D3 x2, D3& b1, D3& b2) {
  const D3 x01 = x1 - x0, x02 = x2 - x0;
  b1 = nrm(x01);
  b2 = nrm(x02 - dot(x02, b1) * b1);
}
__device__ __forceinline__ M2 local2(D3 b1, D3 b2, D3 x01, D3 x02) {
  return m2(dot(b1, x01), dot(b1, x02), dot(b2, x01), dot(b2, x02));
}
__device__ __forceinline__ void elastic_rest(D3 x0, D3 x1, D3 x2, double& W, M2& invDm) {
  const D3 x01 = x1 - x0, x02 = x2 - x0;
  W = 0.5 * len(cross(x01, x02));
  D3 b1, b2;
  frame(x0, x1, x2, b1, b2);
  invDm = inv(local2(b1, b2, x01, x02));
}
__device__ __forceinline__ double elastic_energy(D3 x0, D3 x1, D3 x2, double W, M2 invDm, double mu, double la) {
  D3 b1, b2;
  frame(x0, x1, x2, b1, b2);
  const M2 F = mul(local2(b1, b2, x1 - x0, x2 - x0), invDm);
  const M2 E = scl(0.5, add(mul(tr(F), F), m2(-1.0, 0.0, 0.0, -1.0)));
  const double te = trace(E);
  return W * (mu * trace(mul(E, tr(E))) + 0.5 * la * te * te);
}
__device__ __forceinline__ void elastic_force(D3 x0, D3 x1, D3 x2, double W, M2 invDm, double mu, double la,
                                              double coeff, D3 f[3]) {
  D3 b1, b2;
  frame(x0, x1, x2, b1, b2);
  const M2 F = mul(local2(b1, b2, x1 - x0, x2 - x0), invDm);
  const M2 E = scl(0.5, add(mul(tr(F), F), m2(-1.0, 0.0, 0.0, -1.0)));
  const double te = trace(E);
  const M2 P = mul(F, add(scl(2.0 * mu, E), m2(la * te, 0.0, 0.0, la * te)));
  const M2 H = mul(scl(coeff * W, P), tr(invDm));
  f[1] = H.a * b1 + H.c * b2;
  f[2] = H.b * b1 + H.d * b2;
  f[0] = -f[1] - f[2];
}
// H dx with the frame held fixed (elastic_energy.py:148-220)
__device__ __forceinline__ void elastic_hdx(D3 x0, D3 x1, D3 x2, D3 d0, D3 d1, D3 d2, double W, M2 invDm, double mu,
                                            double la, D3 f[3]) {
  D3 b1, b2;
  frame(x0, x1, x2, b1, b2);
  const M2 F = mul(local2(b1, b2, x1 - x0, x2 - x0), invDm);
  const M2 dF = mul(local2(b1, b2, d1 - d0, d2 - d0), invDm);
  const M2 E = scl(0.5, add(mul(tr(F), F), m2(-1.0, 0.0, 0.0, -1.0)));
  const double te = trace(E);
  const M2 dE = scl(0.5, add(mul(tr(dF), F), mul(tr(F), dF)));
  const double tde = trace(dE);
  const M2 dP = add(mul(dF, add(scl(2.0 * mu, E), m2(la * te, 0.0, 0.0, la * te))),
                    mul(F, add(scl(2.0 * mu, dE), m2(la * tde, 0.0, 0.0, la * tde))));
  const M2 dH = mul(scl(W, dP), tr(invDm));
  f[1] = dH.a * b1 + dH.c * b2;
  f[2] = dH.b * b1 + dH.d * b2;
  f[0] = -f[1] - f[2];
}

// -------------------------------------------------------------- hinges
// 3x3 row-major
struct M3 {
  double m[9];
};
__device__ __forceinline__ M3 outer(D3 a, D3 b) {
  M3 r;
  r.m[0] = a.x * b.x; r.m[1] = a.x * b.y; r.m[2] = a.x * b.z;
  r.m[3] = a.y * b.x; r.m[4] = a.y * b.y; r.m[5] = a.y * b.z;
  r.m[6] = a.z * b.x; r.m[7] = a.z * b.y; r.m[8] = a.z * b.z;
  return r;
}
__device__ __forceinline__ M3 operator+(M3 a, M3 b) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = a.m[i] + b.m[i];
  return r;
}
__device__ __forceinline__ M3 operator-(M3 a, M3 b) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = a.m[i] - b.m[i];
  return r;
}
__device__ __forceinline__ M3 operator*(double s, M3 a) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = s * a.m[i];
  return r;
}
__device__ __forceinline__ M3 T3(M3 a) {
  M3 r;
  r.m[0] = a.m[0]; r.m[1] = a.m[3]; r.m[2] = a.m[6];
  r.m[3] = a.m[1]; r.m[4] = a.m[4]; r.m[5] = a.m[7];
  r.m[6] = a.m[2]; r.m[7] = a.m[5]; r.m[8] = a.m[8];
  return r;
}
__device__ __forceinline__ M3 sym(M3 a) { return a + T3(a); }
__device__ __forceinline__ D3 mv(const double* __restrict__ a, D3 v) {
  return mk3(a[0] * v.x + a[1] * v.y + a[2] * v.z, a[3] * v.x + a[4] * v.y + a[5] * v.z,
             a[6] * v.x + a[7] * v.y + a[8] * v.z);
}

// signed dihedral angle of the flaps (x0; x1, x2) and (x3; x1, x2)
// (hinge_energy.py:4-13)
__device__ __forceinline__ double hinge_angle(D3 x0, D3 x1, D3 x2, D3 x3) {
  const D3 e12 = nrm(x2 - x1);
  const D3 n1 = nrm(cross(e12, x0 - x1));
  const D3 n2 = nrm(cross(x3 - x1, e12));
  return atan2(dot(cross(n1, n2), e12), dot(n1, n2));
}

// Force (coeff * dE/dx) and the 4x4 Hessian blocks of one hinge, stored
// row-major as a 12x12 matrix H[12 * (3i + a) + 3j + b] = blk[i][j](a, b)
// (hinge_energy.py:89-255; Tamstorf and Grinspun 2013, "Discrete bending
// forces and their Jacobians").
__device__ void hinge_diff(D3 x0, D3 x1, D3 x2, D3 x3, double rest, double elen, double k, double coeff, D3 g[4],
                           double* __restrict__ H) {
  const D3 e0 = x2 - x1, e1 = x0 - x2, e2 = x0 - x1, e1_ = x3 - x2, e2_ = x3 - x1;
  const double e0s = len2(e0), e1s = len2(e1), e2s = len2(e2), e1s_ = len2(e1_), e2s_ = len2(e2_);
  const double l0 = sqrt(e0s), l1 = sqrt(e1s), l2 = sqrt(e2s), l1_ = sqrt(e1s_), l2_ = sqrt(e2s_);
  const D3 u0 = nrm(e0), u1 = nrm(e1), u2 = nrm(e2), u1_ = nrm(e1_), u2_ = nrm(e2_);
  const double ca1 = dot(u2, u0), ca2 = -dot(u1, u0), ca1_ = dot(u2_, u0), ca2_ = -dot(u1_, u0);
  const double h0 = len(cross(e1, e2)) / l0, h1 = len(cross(e0, e2)) / l1, h2 = len(cross(e0, e1)) / l2;
  const double h0_ = len(cross(e1_, e2_)) / l0, h1_ = len(cross(e0, e2_)) / l1_, h2_ = len(cross(e0, e1_)) / l2_;
  const D3 n = nrm(cross(e0, e2)), n_ = nrm(cross(e2_, e0));
  const double th = atan2(dot(cross(n, n_), e0) / l0, dot(n, n_));
  const double dE = k * (th - rest) * elen, d2E = k * elen;
  const D3 t1 = n * (ca2 / h1) + n_ * (ca2_ / h1_);
  const D3 t2 = n * (ca1 / h2) + n_ * (ca1_ / h2_);
  const D3 t0 = n * (-1.0 / h0), t3 = n_ * (-1.0 / h0_);
  g[0] = dE * t0 * coeff;
  g[1] = dE * t1 * coeff;
  g[2] = dE * t2 * coeff;
  g[3] = dE * t3 * coeff;
  const D3 m0 = nrm(cross(e0, n)), m1 = nrm(cross(e1, n)), m2v = nrm(cross(n, e2));
  const D3 m0_ = nrm(cross(n_, e0)), m1_ = nrm(cross(n_, e1_)), m2_ = nrm(cross(e2_, n_));
  const M3 M0 = outer(n, m0), M1 = outer(n, m1), M2m = outer(n, m2v);
  const M3 M0_ = outer(n_, m0_), M1_ = outer(n_, m1_), M2_ = outer(n_, m2_);
  const M3 N0 = (1.0 / e0s) * M0, N0_ = (1.0 / e0s) * M0_;
  const M3 P10 = (ca1 / (h1 * h0)) * T3(M0), P11 = (ca1 / (h1 * h1)) * T3(M1), P12 = (ca1 / (h1 * h2)) * T3(M2m);
  const M3 P20 = (ca2 / (h2 * h0)) * T3(M0), P21 = (ca2 / (h2 * h1)) * T3(M1), P22 = (ca2 / (h2 * h2)) * T3(M2m);
  const M3 P10_ = (ca1_ / (h1_ * h0_)) * T3(M0_), P11_ = (ca1_ / (h1_ * h1_)) * T3(M1_),
           P12_ = (ca1_ / (h1_ * h2_)) * T3(M2_);
  const M3 P20_ = (ca2_ / (h2_ * h0_)) * T3(M0_), P21_ = (ca2_ / (h2_ * h1_)) * T3(M1_),
           P22_ = (ca2_ / (h2_ * h2_)) * T3(M2_);
  const M3 Q0 = (1.0 / (h0 * h0)) * M0, Q1 = (1.0 / (h0 * h1)) * M1, Q2 = (1.0 / (h0 * h2)) * M2m;
  const M3 Q0_ = (1.0 / (h0_ * h0_)) * M0_, Q1_ = (1.0 / (h0_ * h1_)) * M1_, Q2_ = (1.0 / (h0_ * h2_)) * M2_;
  M3 b[4][4];
  b[0][0] = (-dE) * sym(Q0) + d2E * outer(t0, t0);
  b[3][3] = (-dE) * sym(Q0_) + d2E * outer(t3, t3);
  b[1][1] = dE * (sym(P11) - N0 + sym(P11_) - N0_) + d2E * outer(t1, t1);
  b[2][2] = dE * (sym(P22) - N0 + sym(P22_) - N0_) + d2E * outer(t2, t2);
  b[1][0] = dE * (P10 - Q1) + d2E * outer(t1, t0);
  b[2][0] = dE * (P20 - Q2) + d2E * outer(t2, t0);
  b[1][3] = dE * (P10_ - Q1_) + d2E * outer(t1, t3);
  b[2][3] = dE * (P20_ - Q2_) + d2E * outer(t2, t3);
  b[1][2] = dE * (P12 + T3(P21) + N0 + P12_ + T3(P21_) + N0_) + d2E * outer(t1, t2);
  b[0][3] = d2E * outer(t0, t3);
  b[0][1] = T3(b[1][0]);
  b[0][2] = T3(b[2][0]);
  b[3][1] = T3(b[1][3]);
  b[3][2] = T3(b[2][3]);
  b[2][1] = T3(b[1][2]);
  b[3][0] = T3(b[0][3]);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) H[12 * (3 * i + a) + 3 * j + c] = b[i][j].m[3 * a + c];
}

// The hinge Hessian is projected on the PSD cone by k3_spd below
// (wp.spd_project_blocks of the authors' Warp fork, utils_kernels.py:20-41,
// which is not in the reference): the leading 9x9 (vertices x0, x1, x2) by
// it_max cyclic Jacobi sweeps (Golub and Van Loan, 8.5.2), negative
// eigenvalues set to zero, then block row and column 3 rebuilt from zero
// block-row sums.

// -------------------------------------------------------------- ACCD
// Additive CCD of one contact along the search direction v
// (ccd_kernels.py:8-110): the largest step (at most 1) that keeps the
// contact's gap above (1 - slack) of its current value.
__device__ double accd_pair(D3 x[4], D3 v[4], int type1, double slack, double thick, int max_iters, double ee_thres) {
  const D3 va = (v[0] + v[1] + v[2] + v[3]) / 4.0;
  D3 w[4];
  for (int k = 0; k < 4; k++) w[k] = v[k] - va;
  const double lv = type1 == BT_PT ? len(w[0]) + fmax(len(w[1]), fmax(len(w[2]), len(w[3])))
                                   : fmax(len(w[0]), len(w[1])) + fmax(len(w[2]), len(w[3]));
  if (lv < 1e-12) return 1.0;
  D3 p[4] = {x[0], x[1], x[2], x[3]};
  auto dist = [&]() {
    if (type1 == BT_PT) return pt_distance(p[0], p[1], p[2], p[3], pt_classify(p[0], p[1], p[2], p[3]));
    return ee_distance(p[0], p[1], p[2], p[3], ee_classify(p[0], p[1], p[2], p[3], ee_thres));
  };
  auto gap = [&](double d) {
    const double ds = d * d;
    const double r = (ds - thick * thick) / (sqrt(ds) + thick);
    return r > 0.0 ? r : 0.0;
  };
  double dmt = gap(dist());
  const double gmin = (1.0 - slack) * dmt;
  double t = 0.0, tl = slack * dmt / lv;
  for (int i = 0;; i++) {
    for (int k = 0; k < 4; k++) p[k] = p[k] + tl * w[k];
    dmt = gap(dist());
    if ((i > 0 && dmt <= gmin) || i >= max_iters) return t;
    t = t + tl;
    if (t > 1.0) return 1.0;
    tl = 0.8 * dmt / lv;
  }
}

// ---------------------------------------------------------- parameters
struct Cfg {
  int n_newton, n_cg, n_ls, spd_iters, ccd_iters;
  double d_hat, kappa, radius, ee_thres, slack, thick, young, poisson, hinge_k, m2gt_k, min_angle;
  long long max_blocks;
  double mu, la;   // Lame parameters of young / poisson (energy.py:353-356)
};

// device scalars (one array of doubles; the u64 slots hold order-preserving
// bits of non-negative doubles for atomicMin)
enum {
  SC_E = 0,        // energy at the current positions
  SC_EPREV = 1,    // energy before the line search
  SC_VAV = 2,      // CG: v . A v
  SC_ZR0 = 3,      // CG: z . r of even / odd iterations
  SC_ZR1 = 4,
  SC_LS = 5,       // line search: halvings taken
  SC_PART = 6,     // energies by part (4)
  SC_MAXABS = 10,  // float bits of the largest |coordinate| (uint)
  SC_CCD = 11,     // u64 bits of the CCD step
  SC_MINANG = 12,  // u64 bits of the smallest corner angle (degrees)
  SC_GMAXABS = 13, // float bits of the target's largest |coordinate| (uint)
  SC_TOTAL = 14,   // u64: 64-bit total of a count array (k3_total64)
  SC_COUNT = 16
};

// corner indices outside [0, nv) counted into err
__global__ void k3_check_idx(const int* __restrict__ idx, long long n, int nv, int* __restrict__ err) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (idx[i] < 0 || idx[i] >= nv)) atomicAdd(err, 1);
}

// --------------------------------------------------------------- CSR
// vertex -> element slots (element * K + corner), sorted per vertex
__global__ void k3_corner_count(const int* __restrict__ idx, long long n, int K, int* __restrict__ cnt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * K) return;
  atomicAdd(&cnt[idx[i]], 1);
}
__global__ void k3_corner_fill(const int* __restrict__ idx, long long n, int K, const int* __restrict__ off,
                               int* __restrict__ cur, int* __restrict__ list) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * K) return;
  const int v = idx[i];
  list[off[v] + atomicAdd(&cur[v], 1)] = (int)i;
}
// Each entry's place in its vertex's segment is its rank there (entries are
// distinct slot numbers): tmp (filled in any order) -> list (ascending).
__global__ void k3_rank_place(const int* __restrict__ idx, long long m, const int* __restrict__ off,
                              const int* __restrict__ tmp, int* __restrict__ list) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int s = tmp[i], v = idx[s], a = off[v], b = off[v + 1];
  int r = 0;
  for (int j = a; j < b; j++) r += tmp[j] < s;
  list[a + r] = s;
}

// ------------------------------------------------------- registration
// Unique edges (a, b), a < b, numbered in ascending (a, b) order: vertex a
// owns its larger neighbours; a neighbour counts at its first appearance in
// a's face list, and its rank is the number of smaller unique larger
// neighbours.
__device__ __forceinline__ bool nb_first(const int* __restrict__ F, const int* __restrict__ vf, int a0, int s, int c,
                                         int b) {
  // b first seen at slot s, corner c (earlier slots, then earlier corners)
  for (int t = a0; t < s; t++) {
    const int f = vf[t] / 3;
    if (F[3 * f] == b || F[3 * f + 1] == b || F[3 * f + 2] == b) return false;
  }
  const int f = vf[s] / 3;
  for (int k = 0; k < c; k++)
    if (F[3 * f + k] == b) return false;
  return true;
}
template <bool EMIT>
__global__ void k3_edges(int nV, const int* __restrict__ F, const int* __restrict__ vf_off, const int* __restrict__ vf,
                         int* __restrict__ ecnt, const int* __restrict__ eoff, int* __restrict__ E) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nV) return;
  const int a0 = vf_off[a], a1 = vf_off[a + 1];
  int n = 0;
  for (int s = a0; s < a1; s++) {
    const int f = vf[s] / 3;
    for (int c = 0; c < 3; c++) {
      const int b = F[3 * f + c];
      if (b <= a || !nb_first(F, vf, a0, s, c, b)) continue;
      if (!EMIT) {
        n++;
        continue;
      }
      int rank = 0;   // unique larger neighbours below b
      for (int s2 = a0; s2 < a1; s2++) {
        const int f2 = vf[s2] / 3;
        for (int c2 = 0; c2 < 3; c2++) {
          const int b2 = F[3 * f2 + c2];
          if (b2 > a && b2 < b && nb_first(F, vf, a0, s2, c2, b2)) rank++;
        }
      }
      E[2 * (eoff[a] + rank)] = a;
      E[2 * (eoff[a] + rank) + 1] = b;
    }
  }
  if (!EMIT) ecnt[a] = n;
}

// One hinge per edge (hinge_energy.py:16-50): the face holding the directed
// edge a -> b (a < b) gives x0, the face holding b -> a gives x3; an edge
// in any other number of faces is an error (the reference asserts one
// hinge per edge, energy.py:626-628).  Rest angle and edge length.
__global__ void k3_hinges(int nE, const int* __restrict__ E, const int* __restrict__ F, const int* __restrict__ vf_off,
                          const int* __restrict__ vf, const double* __restrict__ X, int4* __restrict__ hinge,
                          double2* __restrict__ rest, int* __restrict__ err) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  const int a = E[2 * e], b = E[2 * e + 1];
  int nf = 0, nb = 0, i0 = -1, i3 = -1;
  for (int s = vf_off[a]; s < vf_off[a + 1]; s++) {
    const int f = vf[s] / 3, k = vf[s] % 3;
    if (F[3 * f + (k + 1) % 3] == b) { nf++; i0 = F[3 * f + (k + 2) % 3]; }
    if (F[3 * f + (k + 2) % 3] == b) { nb++; i3 = F[3 * f + (k + 1) % 3]; }
  }
  if (nf != 1 || nb != 1) {
    atomicAdd(err, 1);
    i0 = i3 = a;
  }
  hinge[e] = make_int4(i0, a, b, i3);
  const D3 p0 = ld3(X, i0), p1 = ld3(X, a), p2 = ld3(X, b), p3 = ld3(X, i3);
  rest[e] = make_double2(nf == 1 && nb == 1 ? hinge_angle(p0, p1, p2, p3) : 0.0, len(p1 - p2));
}

// Mixed Voronoi area per vertex (Meyer, Desbrun, Schroeder and Barr 2003),
// the diagonal of igl::massmatrix(V, F, MASSMATRIX_TYPE_VORONOI)
// (system.py:293-295), gathered over the vertex's faces in ascending order.
__global__ void k3_voronoi(int nV, const double* __restrict__ X, const int* __restrict__ F,
                           const int* __restrict__ vf_off, const int* __restrict__ vf, double* __restrict__ vor) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  double acc = 0.0;
  for (int s = vf_off[v]; s < vf_off[v + 1]; s++) {
    const int f = vf[s] / 3, i = vf[s] % 3;
    D3 p[3];
    for (int k = 0; k < 3; k++) p[k] = ld3(X, F[3 * f + k]);
    const double A = 0.5 * len(cross(p[1] - p[0], p[2] - p[0]));
    double c[3], cot[3];
    for (int k = 0; k < 3; k++) {
      const D3 a = p[(k + 1) % 3] - p[k], b = p[(k + 2) % 3] - p[k];
      c[k] = dot(a, b);
      cot[k] = c[k] / len(cross(a, b));
    }
    if (c[0] < 0.0 || c[1] < 0.0 || c[2] < 0.0) {
      acc += c[i] < 0.0 ? 0.5 * A : 0.25 * A;
      continue;
    }
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    acc += (len2(p[j] - p[i]) * cot[k] + len2(p[k] - p[i]) * cot[j]) / 8.0;
  }
  vor[v] = acc;
}

// smallest corner angle in degrees (igl.internal_angles, system.py:288-292)
// and the membrane rest state of every face (elastic_energy.py:7-37)
__global__ void k3_faces_rest(int nF, const double* __restrict__ X, const int* __restrict__ F,
                              double* __restrict__ W, double4* __restrict__ invDm, unsigned long long* minang) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  D3 p[3];
  for (int k = 0; k < 3; k++) p[k] = ld3(X, F[3 * f + k]);
  double m = 180.0;
  for (int i = 0; i < 3; i++) {
    const D3 a = p[(i + 1) % 3] - p[i], b = p[(i + 2) % 3] - p[i];
    double c = dot(a, b) / (len(a) * len(b));
    if (c != c) continue;
    c = fmax(-1.0, fmin(1.0, c));
    m = fmin(m, acos(c) * 180.0 / 3.14159265358979323846);
  }
  atomicMin(minang, (unsigned long long)__double_as_longlong(m));
  double w;
  M2 id;
  elastic_rest(p[0], p[1], p[2], w, id);
  W[f] = w;
  invDm[f] = make_double4(id.a, id.b, id.c, id.d);
}

// --------------------------------------------------------- float trees
// Positions rounded to float for the LBVH boxes, and the largest
// |coordinate| (its float bits, for the boxes' rounding margin).
// (a grid-stride loop over at most S3_TF_BLOCKS workgroups of 256 lanes,
// one atomic per workgroup: the atomics all go to one word)
constexpr int S3_TF_BLOCKS = 1024;
__global__ void __launch_bounds__(256) k3_to_float(long long n, const double* __restrict__ X, float* __restrict__ P,
                                                   unsigned* __restrict__ maxabs) {
  __shared__ float red[4];
  float m = 0.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = (float)X[i];
    P[i] = v;
    float a = fabsf(v);
    if (!(a <= 3.0e38f)) a = 3.0e38f;
    m = fmaxf(m, a);
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(maxabs, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}
// the 64-bit total of n int counts, added into *out (a grid-stride loop
// over at most S3_TF_BLOCKS workgroups of 256 lanes, one atomic each): the
// int scans of the counts are taken only once this total is known to fit
__global__ void __launch_bounds__(256) k3_total64(const int* __restrict__ cnt, long long n,
                                                  unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4];
  unsigned long long t = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    t += (unsigned long long)(unsigned)max(cnt[i], 0);
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}
__global__ void k3_iota(int n, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
// an edge (a, b) as the triangle (a, b, b): its box is the segment's
__global__ void k3_edge_tris(int nE, const int* __restrict__ E, int* __restrict__ T) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  T[3 * e] = E[2 * e];
  T[3 * e + 1] = E[2 * e + 1];
  T[3 * e + 2] = E[2 * e + 1];
}
// The float boxes hold the float-rounded corners, within half an ulp of
// the float64 ones; margin = 4 ulps at the largest coordinate covers that.
__device__ __forceinline__ float box_margin(const double* sc, int slot) {
  const float m = __uint_as_float(((const unsigned*)sc)[2 * slot]);
  return m * 4.0f * 5.9604645e-8f + 1e-30f;
}
__device__ __forceinline__ bool box_hit(float4 A, float4 B, const float lo[3], const float hi[3]) {
  return !(hi[0] < A.x || B.x < lo[0]) && !(hi[1] < A.y || B.y < lo[1]) && !(hi[2] < A.z || B.z < lo[2]);
}

// ----------------------------------------------------------- targets
// Nearest point of the target surface for every vertex: a depth-first walk
// of the target's LBVH (ropes) pruning nodes whose box lies farther than
// the best squared distance found (boxes grown by the rounding margin);
// equal distances keep the lowest triangle index, as a scan of the
// triangles in order with a strict < does.
__global__ void k3_nearest(int nV, const double* __restrict__ q, const double* __restrict__ gV,
                           const int* __restrict__ gF, const float4* __restrict__ nodeA,
                           const float4* __restrict__ nodeB, int n_nodes, const double* __restrict__ sc,
                           double* __restrict__ target, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const D3 x = ld3(q, i);
  const double mg = (double)box_margin(sc, SC_GMAXABS);
  auto lbound = [&](float4 A, float4 B) {
    const double dx = fmax(fmax((double)A.x - mg - x.x, x.x - ((double)B.x + mg)), 0.0);
    const double dy = fmax(fmax((double)A.y - mg - x.y, x.y - ((double)B.y + mg)), 0.0);
    const double dz = fmax(fmax((double)A.z - mg - x.z, x.z - ((double)B.z + mg)), 0.0);
    return dx * dx + dy * dy + dz * dz;
  };
  auto leaf = [&](int t) {
    return closest_on_tri(x, ld3(gV, gF[3 * t]), ld3(gV, gF[3 * t + 1]), ld3(gV, gF[3 * t + 2]));
  };
  // a first bound: descend into the nearer child down to one leaf (the right
  // child of a node is the escape of its left child)
  int node = 0, t0 = -1;
  for (int depth = 0; depth <= n_nodes; depth++) {
    const int c = __float_as_int(nodeB[node].w);
    if (c < 0) {
      t0 = ~c;
      break;
    }
    const float4 Al = nodeA[c], Bl = nodeB[c];
    const int r = __float_as_int(Al.w);
    node = lbound(Al, Bl) <= lbound(nodeA[r], nodeB[r]) ? c : r;
  }
  double best = INFINITY;
  int bestf = -1;
  D3 bp = x;
  if (t0 >= 0) {
    bp = leaf(t0);
    best = len2(bp - x);
    bestf = t0;
  }
  node = 0;
  long long steps = 0;
  while (node != -1) {
    if (++steps > 2LL * n_nodes + 8) {
      atomicAdd(err, 1);
      break;
    }
    const float4 A = nodeA[node], B = nodeB[node];
    const int esc = __float_as_int(A.w);
    if (lbound(A, B) <= best) {
      const int c = __float_as_int(B.w);
      if (c >= 0) {
        node = c;
        continue;
      }
      const int t = ~c;
      const D3 cp = leaf(t);
      const double d = len2(cp - x);
      if (d < best || (d == best && t < bestf)) {
        best = d;
        bestf = t;
        bp = cp;
      }
    }
    node = esc;
  }
  st3(target, i, bp);
}

// ----------------------------------------------------------- contacts
// one query's pairs [base, base + n) sorted by key ce1 (insertion sort)
__device__ void sort_run(int base, int n, int4* __restrict__ cidx, int2* __restrict__ ctyp, double* __restrict__ cd,
                         int* __restrict__ ce1) {
  for (int a = base + 1; a < base + n; a++) {
    const int4 k = cidx[a];
    const int2 kt = ctyp[a];
    const double kd = cd[a];
    const int ke = ce1[a];
    int b = a - 1;
    while (b >= base && ce1[b] > ke) {
      cidx[b + 1] = cidx[b];
      ctyp[b + 1] = ctyp[b];
      cd[b + 1] = cd[b];
      ce1[b + 1] = ce1[b];
      b--;
    }
    cidx[b + 1] = k;
    ctyp[b + 1] = kt;
    cd[b + 1] = kd;
    ce1[b + 1] = ke;
  }
}

// Contact search through LBVHs (meshes above the exhaustive search's
// limit).  Point-triangle pairs of vertex i0 within thr (contact_detection.py:
// 120-216): the LBVH of the faces is walked with i0's box grown by thr and
// the rounding margin; faces with i0 as a corner are skipped.  Count pass
// (off == nullptr) writes cnt[i0]; emit pass writes the pairs at off[i0],
// then sorts them by triangle (kept in ce1).
__global__ void k3_pt_query(int nV, const double* __restrict__ q, const int* __restrict__ F,
                            const float4* __restrict__ nodeA, const float4* __restrict__ nodeB, int n_nodes,
                            double thr, const double* __restrict__ sc, int* __restrict__ cnt,
                            const int* __restrict__ off, int4* __restrict__ cidx, int2* __restrict__ ctyp,
                            double* __restrict__ cd, int* __restrict__ ce1, int* __restrict__ err) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 >= nV) return;
  const D3 x0 = ld3(q, i0);
  const double pad = thr + (double)box_margin(sc, SC_MAXABS);
  const float lo[3] = {__double2float_rd(x0.x - pad), __double2float_rd(x0.y - pad), __double2float_rd(x0.z - pad)};
  const float hi[3] = {__double2float_ru(x0.x + pad), __double2float_ru(x0.y + pad), __double2float_ru(x0.z + pad)};
  const int base = off ? off[i0] : 0;
  int n = 0, node = 0;
  long long steps = 0;
  while (node != -1) {
    if (++steps > 2LL * n_nodes + 8) {
      atomicAdd(err, 1);
      break;
    }
    const float4 A = nodeA[node], B = nodeB[node];
    const int esc = __float_as_int(A.w);
    if (box_hit(A, B, lo, hi)) {
      const int c = __float_as_int(B.w);
      if (c >= 0) {
        node = c;
        continue;
      }
      const int t = ~c;
      const int i1 = F[3 * t], i2 = F[3 * t + 1], i3 = F[3 * t + 2];
      if (i0 != i1 && i0 != i2 && i0 != i3) {
        const D3 x1 = ld3(q, i1), x2 = ld3(q, i2), x3 = ld3(q, i3);
        const int ty = pt_classify(x0, x1, x2, x3);
        const double d = pt_distance(x0, x1, x2, x3, ty);
        if (d <= thr) {
          if (off) {
            cidx[base + n] = make_int4(i0, i1, i2, i3);
            ctyp[base + n] = make_int2(BT_PT, ty);
            cd[base + n] = d;
            ce1[base + n] = t;
          }
          n++;
        }
      }
    }
    node = esc;
  }
  if (!off) {
    cnt[i0] = n;
    return;
  }
  sort_run(base, n, cidx, ctyp, cd, ce1);
}

// Edge-edge pairs of edge e0 with the later edges e1 > e0 that share no
// vertex, within thr; the LBVH of the edges is walked with e0's box grown
// by thr and the margin.  Counts at cnt[nV + e0]; pairs sorted by e1.
__global__ void k3_ee_query(int nE, int nV, const double* __restrict__ q, const int* __restrict__ E,
                            const float4* __restrict__ nodeA, const float4* __restrict__ nodeB, int n_nodes,
                            double thr, double ee_thres, const double* __restrict__ sc, int* __restrict__ cnt,
                            const int* __restrict__ off, int4* __restrict__ cidx, int2* __restrict__ ctyp,
                            double* __restrict__ cd, int* __restrict__ ce1, int* __restrict__ err) {
  const int e0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (e0 >= nE) return;
  const int i0 = E[2 * e0], i1 = E[2 * e0 + 1];
  const D3 x0 = ld3(q, i0), x1 = ld3(q, i1);
  const double pad = thr + (double)box_margin(sc, SC_MAXABS);
  const float lo[3] = {__double2float_rd(fmin(x0.x, x1.x) - pad), __double2float_rd(fmin(x0.y, x1.y) - pad),
                       __double2float_rd(fmin(x0.z, x1.z) - pad)};
  const float hi[3] = {__double2float_ru(fmax(x0.x, x1.x) + pad), __double2float_ru(fmax(x0.y, x1.y) + pad),
                       __double2float_ru(fmax(x0.z, x1.z) + pad)};
  const int base = off ? off[nV + e0] : 0;
  int n = 0, node = 0;
  long long steps = 0;
  while (node != -1) {
    if (++steps > 2LL * n_nodes + 8) {
      atomicAdd(err, 1);
      break;
    }
    const float4 A = nodeA[node], B = nodeB[node];
    const int esc = __float_as_int(A.w);
    if (box_hit(A, B, lo, hi)) {
      const int c = __float_as_int(B.w);
      if (c >= 0) {
        node = c;
        continue;
      }
      const int e1 = ~c;
      if (e1 > e0) {
        const int i2 = E[2 * e1], i3 = E[2 * e1 + 1];
        if (i0 != i2 && i0 != i3 && i1 != i2 && i1 != i3) {
          const D3 x2 = ld3(q, i2), x3 = ld3(q, i3);
          const int ty = ee_classify(x0, x1, x2, x3, ee_thres);
          const double d = ee_distance(x0, x1, x2, x3, ty);
          if (d <= thr) {
            if (off) {
              cidx[base + n] = make_int4(i0, i1, i2, i3);
              ctyp[base + n] = make_int2(BT_EE, ty);
              cd[base + n] = d;
              ce1[base + n] = e1;
            }
            n++;
          }
        }
      }
    }
    node = esc;
  }
  if (!off) {
    cnt[nV + e0] = n;
    return;
  }
  sort_run(base, n, cidx, ctyp, cd, ce1);
}

// Exhaustive contact search (meshes up to the limit), in three passes:
//  1. box filter: every query against every candidate, the candidates split
//     into P ranges (part = blockIdx.y) and staged through LDS in tiles of
//     float boxes rounded outward from the float64 corners (no margin
//     needed); counted, then emitted as (query, candidate) pairs at scanned
//     offsets laid out query-major (q P + part), so each query's pairs are
//     in ascending candidate order;
//  2. narrow phase: one lane per pair (no lane waits on another's
//     classification);
//  3. compaction of the pairs within thr, keeping that order.
constexpr int S3_TILE = 256;
__global__ void k3_boxes(int nF, int nE, const double* __restrict__ q, const int* __restrict__ F,
                         const int* __restrict__ E, float4* __restrict__ flo, float4* __restrict__ fhi,
                         float4* __restrict__ elo, float4* __restrict__ ehi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nF + nE) return;
  D3 a, b, c;
  if (i < nF) {
    a = ld3(q, F[3 * i]); b = ld3(q, F[3 * i + 1]); c = ld3(q, F[3 * i + 2]);
  } else {
    a = ld3(q, E[2 * (i - nF)]); b = ld3(q, E[2 * (i - nF) + 1]); c = b;
  }
  const float4 lo = make_float4(__double2float_rd(fmin(a.x, fmin(b.x, c.x))), __double2float_rd(fmin(a.y, fmin(b.y, c.y))),
                                __double2float_rd(fmin(a.z, fmin(b.z, c.z))), 0.0f);
  const float4 hi = make_float4(__double2float_ru(fmax(a.x, fmax(b.x, c.x))), __double2float_ru(fmax(a.y, fmax(b.y, c.y))),
                                __double2float_ru(fmax(a.z, fmax(b.z, c.z))), 0.0f);
  if (i < nF) {
    flo[i] = lo;
    fhi[i] = hi;
  } else {
    elo[i - nF] = lo;
    ehi[i - nF] = hi;
  }
}
__device__ __forceinline__ bool box_meets(float4 lo, float4 hi, const float qlo[3], const float qhi[3]) {
  return !(qhi[0] < lo.x || hi.x < qlo[0]) && !(qhi[1] < lo.y || hi.y < qlo[1]) && !(qhi[2] < lo.z || hi.z < qlo[2]);
}
// Pass 1.  EE (kind 1): query e0 against edges e1 > e0 sharing no vertex;
// PT (kind 0): query vertex i0 against faces without i0.
template <int KIND, bool EMIT>
__global__ void __launch_bounds__(S3_TILE) k3_box_pairs(int nq, int ncand, int chunk, int P, const double* __restrict__ q,
                                                       const int* __restrict__ F, const int* __restrict__ E,
                                                       const float4* __restrict__ clo, const float4* __restrict__ chi,
                                                       double thr, int* __restrict__ cnt, const int* __restrict__ off,
                                                       int2* __restrict__ pairs) {
  __shared__ float4 slo[S3_TILE], shi[S3_TILE];
  const int part = blockIdx.y;
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = qi < nq;
  int i0 = 0, i1 = 0;
  D3 a = mk3(0.0, 0.0, 0.0), b = a;
  if (live) {
    if (KIND == 0) {
      i0 = qi;
      a = b = ld3(q, i0);
    } else {
      i0 = E[2 * qi];
      i1 = E[2 * qi + 1];
      a = ld3(q, i0);
      b = ld3(q, i1);
    }
  }
  const float qlo[3] = {__double2float_rd(fmin(a.x, b.x) - thr), __double2float_rd(fmin(a.y, b.y) - thr),
                        __double2float_rd(fmin(a.z, b.z) - thr)};
  const float qhi[3] = {__double2float_ru(fmax(a.x, b.x) + thr), __double2float_ru(fmax(a.y, b.y) + thr),
                        __double2float_ru(fmax(a.z, b.z) + thr)};
  const int t0 = part * chunk, t1 = min(ncand, t0 + chunk);
  const int base = EMIT && live ? off[(long long)qi * P + part] : 0;
  int n = 0;
  for (int tb = t0; tb < t1; tb += S3_TILE) {
    const int m = min(S3_TILE, t1 - tb);
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      slo[j] = clo[tb + j];
      shi[j] = chi[tb + j];
    }
    __syncthreads();
    if (!live) continue;
    // four box tests per step (independent LDS reads), then their hits in order
    for (int j = KIND == 1 ? max(0, qi + 1 - tb) : 0; j < m; j += 4) {
      bool h[4];
#pragma unroll
      for (int u = 0; u < 4; u++) h[u] = j + u < m && box_meets(slo[j + u], shi[j + u], qlo, qhi);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (!h[u]) continue;
        const int t = tb + j + u;
        if (KIND == 0) {
          if (F[3 * t] == i0 || F[3 * t + 1] == i0 || F[3 * t + 2] == i0) continue;
        } else {
          const int i2 = E[2 * t], i3 = E[2 * t + 1];
          if (i0 == i2 || i0 == i3 || i1 == i2 || i1 == i3) continue;
        }
        if (EMIT) pairs[base + n] = make_int2(qi, t);
        n++;
      }
    }
  }
  if (!EMIT && live) cnt[(long long)qi * P + part] = n;
}
// Pass 2: pairs [0, npt) are point-triangle, [npt, n) edge-edge
__device__ __forceinline__ int4 pair_corners(long long k, long long npt, int2 pr, const int* __restrict__ F,
                                             const int* __restrict__ E) {
  if (k < npt) return make_int4(pr.x, F[3 * pr.y], F[3 * pr.y + 1], F[3 * pr.y + 2]);
  return make_int4(E[2 * pr.x], E[2 * pr.x + 1], E[2 * pr.y], E[2 * pr.y + 1]);
}
__global__ void k3_narrow(long long n, long long npt, const int2* __restrict__ pairs, const double* __restrict__ q,
                          const int* __restrict__ F, const int* __restrict__ E, double thr, double ee_thres,
                          int* __restrict__ hit, int* __restrict__ ptyp, double* __restrict__ pd) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int4 c = pair_corners(k, npt, pairs[k], F, E);
  const D3 x0 = ld3(q, c.x), x1 = ld3(q, c.y), x2 = ld3(q, c.z), x3 = ld3(q, c.w);
  int ty;
  double d;
  if (k < npt) {
    ty = pt_classify(x0, x1, x2, x3);
    d = pt_distance(x0, x1, x2, x3, ty);
  } else {
    ty = ee_classify(x0, x1, x2, x3, ee_thres);
    d = ee_distance(x0, x1, x2, x3, ty);
  }
  hit[k] = d <= thr ? 1 : 0;
  ptyp[k] = ty;
  pd[k] = d;
}
// Pass 3
__global__ void k3_compact(long long n, long long npt, const int2* __restrict__ pairs, const int* __restrict__ F,
                           const int* __restrict__ E, const int* __restrict__ hit, const int* __restrict__ pos,
                           const int* __restrict__ ptyp, const double* __restrict__ pd, int4* __restrict__ cidx,
                           int2* __restrict__ ctyp, double* __restrict__ cd) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !hit[k]) return;
  const int o = pos[k];
  cidx[o] = pair_corners(k, npt, pairs[k], F, E);
  ctyp[o] = make_int2(k < npt ? BT_PT : BT_EE, ptyp[k]);
  cd[o] = pd[k];
}

// ------------------------------------------------------------ energy
// Per-element energies at x into en[] (vertices, faces, hinges, contacts,
// system.py:343-358); contacts are re-classified at x and keep the new type
// and distance (the derivatives use them).
// Line-search step n_ls > 0 after the energy dropped below the energy
// before the search: the positions did not move (k3_ls_pos), so the energy
// and the contacts' types and distances would come out the same, and the
// energy kernels return at once.  n_ls < 0: any other evaluation.
__device__ __forceinline__ bool ls_settled(const double* __restrict__ sc, int n_ls) {
  return n_ls > 0 && sc[SC_E] < sc[SC_EPREV];
}
__global__ void k3_energy_elem(int nV, int nF, int nH, const int* __restrict__ nCp, int n_ls,
                               const double* __restrict__ sc, const double* __restrict__ x,
                               const double* __restrict__ target, const double* __restrict__ vor,
                               const int* __restrict__ F, const double* __restrict__ W,
                               const double4* __restrict__ invDm, const int4* __restrict__ hinge,
                               const double2* __restrict__ rest, const int4* __restrict__ cidx,
                               int2* __restrict__ ctyp, double* __restrict__ cd, Cfg c, double* __restrict__ en) {
  if (ls_settled(sc, n_ls)) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nV + nF + nH + *nCp;
  if (i >= n) return;
  if (i < nV) {
    en[i] = 0.5 * len2(ld3(x, i) - ld3(target, i)) * (vor[i] * c.m2gt_k);
  } else if (i < nV + nF) {
    const int f = (int)(i - nV);
    const double4 m = invDm[f];
    en[i] = elastic_energy(ld3(x, F[3 * f]), ld3(x, F[3 * f + 1]), ld3(x, F[3 * f + 2]), W[f],
                           m2(m.x, m.y, m.z, m.w), c.mu, c.la);
  } else if (i < (long long)nV + nF + nH) {
    const int h = (int)(i - nV - nF);
    const int4 k = hinge[h];
    const double th = hinge_angle(ld3(x, k.x), ld3(x, k.y), ld3(x, k.z), ld3(x, k.w));
    const double dt = th - rest[h].x;
    en[i] = 0.5 * c.hinge_k * dt * dt * rest[h].y;
  } else {
    const int j = (int)(i - nV - nF - nH);
    const int4 k = cidx[j];
    const D3 a = ld3(x, k.x), b = ld3(x, k.y), cc = ld3(x, k.z), d = ld3(x, k.w);
    int t;
    double dist;
    if (ctyp[j].x == BT_PT) {
      t = pt_classify(a, b, cc, d);
      dist = pt_distance(a, b, cc, d, t);
    } else {
      t = ee_classify(a, b, cc, d, c.ee_thres);
      dist = ee_distance(a, b, cc, d, t);
    }
    ctyp[j].y = t;
    cd[j] = dist;
    en[i] = dist < c.d_hat ? bar(dist, c.d_hat) * c.kappa : 0.0;
  }
}

// Deterministic sums by one workgroup of S3_RED lanes: lane t adds elements
// t, t + S3_RED, ... in order, then a fixed tree: a butterfly within each
// 64-lane wave, the waves' sums through LDS, a butterfly over them; every
// lane gets the sum.
constexpr int S3_RED = 1024;
__device__ double block_sum(const double* __restrict__ a, long long n, double* lds) {
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += S3_RED) s += a[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = s;
  __syncthreads();
  // every aligned group of S3_RED / 64 lanes reduces the same wave sums in
  // the same order, so every lane returns the same bits
  double t = lds[lane & (S3_RED / 64 - 1)];
#pragma unroll
  for (int o = S3_RED / 128; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  __syncthreads();   // every lane has read lds before the next call writes it
  return t;
}
// the four parts of en[] and their total: parts summed, then
// ((m2gt + membrane) + hinge) + contact
__global__ void __launch_bounds__(S3_RED) k3_energy_sum(const double* __restrict__ en, int nV, int nF, int nH,
                                                        const int* __restrict__ nCp, int n_ls,
                                                        double* __restrict__ sc, int slot) {
  __shared__ double lds[S3_RED];
  if (ls_settled(sc, n_ls)) return;
  const int nC = *nCp;
  const long long o1 = nV, o2 = o1 + nF, o3 = o2 + nH;
  const double e0 = block_sum(en, nV, lds);
  const double e1 = block_sum(en + o1, nF, lds);
  const double e2 = block_sum(en + o2, nH, lds);
  const double e3 = block_sum(en + o3, nC, lds);
  if (threadIdx.x == 0) {
    sc[slot] = e0 + e1 + e2 + e3;
    sc[SC_PART] = e0;
    sc[SC_PART + 1] = e1;
    sc[SC_PART + 2] = e2;
    sc[SC_PART + 3] = e3;
  }
}
__global__ void __launch_bounds__(S3_RED) k3_sum(const double* __restrict__ a, long long n, double* __restrict__ out) {
  __shared__ double lds[S3_RED];
  const double s = block_sum(a, n, lds);
  if (threadIdx.x == 0) *out = s;
}

// -------------------------------------------------------- derivatives
// Per-element contributions (coeff = -1: forces) into slots:
//   faces     fs[9 f + 3 k ..]          membrane force on corner k
//   hinges    hs[12 h + 3 k ..]         bending force on corner k
//             hd[12 h + 3 k ..]         diagonal of the unprojected block (k, k)
//             hb[144 h ..]              projected 12x12 Hessian
//   contacts  cdd[12 c ..]              distance gradient (for H dx)
//             cg[c] = (g1, g2)          barrier slope * kappa * coeff, curvature * kappa
//             cs[12 c + 3 j ..]         dd_j * g1
//             cq[12 c + 3 j ..]         g2 * dd_j * dd_j (diagonal)
__global__ void k3_diff_faces(int nF, const double* __restrict__ x, const int* __restrict__ F,
                              const double* __restrict__ W, const double4* __restrict__ invDm, Cfg c,
                              double* __restrict__ fs) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  const double4 m = invDm[f];
  D3 o[3];
  elastic_force(ld3(x, F[3 * f]), ld3(x, F[3 * f + 1]), ld3(x, F[3 * f + 2]), W[f], m2(m.x, m.y, m.z, m.w), c.mu,
                c.la, -1.0, o);
  for (int k = 0; k < 3; k++) st3(fs, 3LL * f + k, o[k]);
}
constexpr int S3_HINGE_BLOCK = 64;
__global__ void __launch_bounds__(S3_HINGE_BLOCK) k3_diff_hinges(int nH, const double* __restrict__ x,
                                                                 const int4* __restrict__ hinge,
                                                                 const double2* __restrict__ rest, Cfg c,
                                                                 double* __restrict__ hs, double* __restrict__ hd,
                                                                 double* __restrict__ hb) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nH) return;
  const int4 k = hinge[h];
  D3 g[4];
  double* H = hb + 144LL * h;
  hinge_diff(ld3(x, k.x), ld3(x, k.y), ld3(x, k.z), ld3(x, k.w), rest[h].x, rest[h].y, c.hinge_k, -1.0, g, H);
  for (int j = 0; j < 4; j++) {
    st3(hs, 4LL * h + j, g[j]);
    st3(hd, 4LL * h + j, mk3(H[12 * (3 * j) + 3 * j], H[12 * (3 * j + 1) + 3 * j + 1], H[12 * (3 * j + 2) + 3 * j + 2]));
  }
}

// The projection of every hinge's Hessian by 9 lanes (7 hinges per 64-lane
// workgroup): lane k owns row k of the leading 9x9 in the row updates and
// column k in the column updates of each Jacobi rotation; every lane of a
// hinge computes the rotation from the same entries.  Each update touches
// one row (or column) per lane, as in the serial loops, so the results are
// those of spd_project.  The rebuilt block row and column 3 and the
// reconstruction run on the LDS copy, then the 12x12 goes back.
constexpr int S3_SPD_H = 7;
__global__ void __launch_bounds__(64) k3_spd(int nH, int it_max, double* __restrict__ hb) {
  __shared__ double A[S3_SPD_H][81], Q[S3_SPD_H][81], Hl[S3_SPD_H][144];
  const int w = threadIdx.x / 9, k = threadIdx.x % 9;
  const int h = blockIdx.x * S3_SPD_H + w;
  const bool live = w < S3_SPD_H && h < nH;
  double* A_ = A[live ? w : 0];
  double* Q_ = Q[live ? w : 0];
  double* H_ = Hl[live ? w : 0];
  if (live) {
    const double* H = hb + 144LL * h;
    for (int j = 0; j < 16; j++) H_[16 * k + j] = H[16 * k + j];
  }
  __syncthreads();
  if (live)
    for (int j = 0; j < 9; j++) {
      A_[9 * k + j] = H_[12 * k + j];
      Q_[9 * k + j] = k == j ? 1.0 : 0.0;
    }
  __syncthreads();
  for (int sweep = 0; sweep < it_max; sweep++)
    for (int pp = 0; pp < 8; pp++)
      for (int qq = pp + 1; qq < 9; qq++) {
        double c = 1.0, sn = 0.0;
        bool rot = false;
        if (live) {
          const double apq = A_[9 * pp + qq];
          rot = apq != 0.0;
          if (rot) {
            const double tau = (A_[9 * qq + qq] - A_[9 * pp + pp]) / (2.0 * apq);
            const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = t * c;
          }
        }
        __syncthreads();
        if (rot) {   // A <- A J and Q <- Q J: row k
          const double akp = A_[9 * k + pp], akq = A_[9 * k + qq];
          A_[9 * k + pp] = c * akp - sn * akq;
          A_[9 * k + qq] = sn * akp + c * akq;
          const double qkp = Q_[9 * k + pp], qkq = Q_[9 * k + qq];
          Q_[9 * k + pp] = c * qkp - sn * qkq;
          Q_[9 * k + qq] = sn * qkp + c * qkq;
        }
        __syncthreads();
        if (rot) {   // A <- J^T A: column k
          const double apk = A_[9 * pp + k], aqk = A_[9 * qq + k];
          A_[9 * pp + k] = c * apk - sn * aqk;
          A_[9 * qq + k] = sn * apk + c * aqk;
        }
        __syncthreads();
      }
  if (live) {
    double lam[9];
    for (int i = 0; i < 9; i++) lam[i] = A_[9 * i + i] > 0.0 ? A_[9 * i + i] : 0.0;
    for (int j = 0; j < 9; j++) {
      double sum = 0.0;
      for (int m = 0; m < 9; m++) sum += Q_[9 * k + m] * lam[m] * Q_[9 * j + m];
      H_[12 * k + j] = sum;
    }
  }
  __syncthreads();
  if (live)   // block row 3, column k
    for (int a = 0; a < 3; a++) {
      const double sum = H_[12 * (0 + a) + k] + H_[12 * (3 + a) + k] + H_[12 * (6 + a) + k];
      H_[12 * (9 + a) + k] = -1.0 * sum;
    }
  __syncthreads();
  if (live)   // block column 3, row k: the transpose of block row 3
    for (int cc = 0; cc < 3; cc++) H_[12 * k + 9 + cc] = H_[12 * (9 + cc) + k];
  __syncthreads();
  if (live) {   // block (3, 3), entry k
    const int a = k / 3, cc = k % 3;
    const double sum = H_[12 * a + 9 + cc] + H_[12 * (3 + a) + 9 + cc] + H_[12 * (6 + a) + 9 + cc];
    H_[12 * (9 + a) + 9 + cc] = -1.0 * sum;
  }
  __syncthreads();
  if (live) {
    double* H = hb + 144LL * h;
    for (int j = 0; j < 16; j++) H[16 * k + j] = H_[16 * k + j];
  }
}
__global__ void k3_diff_contacts(const int* __restrict__ nCp, const double* __restrict__ x, const int4* __restrict__ cidx,
                                 const int2* __restrict__ ctyp, const double* __restrict__ cd, Cfg c,
                                 double* __restrict__ cdd, double2* __restrict__ cg, double* __restrict__ cs,
                                 double* __restrict__ cq) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *nCp) return;
  D3 dd[4];
  for (int k = 0; k < 4; k++) dd[k] = mk3(0.0, 0.0, 0.0);
  const double d = cd[j];
  double g1 = 0.0, g2 = 0.0;
  const bool on = d < c.d_hat;
  if (on) {
    const int4 k = cidx[j];
    const D3 a = ld3(x, k.x), b = ld3(x, k.y), cc = ld3(x, k.z), e = ld3(x, k.w);
    if (ctyp[j].x == BT_PT) pt_grad(a, b, cc, e, ctyp[j].y, dd);
    else ee_grad(a, b, cc, e, ctyp[j].y, dd);
    g1 = bar1(d, c.d_hat) * c.kappa * -1.0;
    g2 = bar2(d, c.d_hat) * c.kappa;
  }
  cg[j] = make_double2(g1, g2);
  for (int k = 0; k < 4; k++) {
    st3(cdd, 4LL * j + k, dd[k]);
    st3(cs, 4LL * j + k, on ? dd[k] * g1 : mk3(0.0, 0.0, 0.0));
    st3(cq, 4LL * j + k, on ? g2 * cmul(dd[k], dd[k]) : mk3(0.0, 0.0, 0.0));
  }
}
// Per-vertex gathers by groups of S3_VG lanes over the vertex's slot list:
// its faces, then its hinges, then its contacts (three CSR segments read as
// one list).  Lane g adds entries g, g + S3_VG, g + 2 S3_VG, ... in order
// (its list loads, then its slot loads, S3_VB at a time in flight), then
// the lanes' partial sums meet in a butterfly: a fixed tree for a given list
// length, so the sum is the same on every run.  Faces add no diagonal term.
constexpr int S3_VG = 32, S3_VB = 8;
struct VSeg {
  const int* list;
  int a, n;
  const double* s1;
  const double* s2;
};
template <bool TWO>
__device__ __forceinline__ void gather3(const VSeg sf, const VSeg sh, const VSeg sc_, int g, D3& sum1, D3& sum2) {
  const int tot = sf.n + sh.n + sc_.n;
  D3 acc1 = mk3(0.0, 0.0, 0.0), acc2 = acc1;
  for (int b0 = 0; b0 < tot; b0 += S3_VG * S3_VB) {
    int id[S3_VB];
    const double* p1[S3_VB];
    const double* p2[S3_VB];
#pragma unroll
    for (int u = 0; u < S3_VB; u++) {
      const int j = b0 + u * S3_VG + g;
      const bool f = j < sf.n, h = !f && j < sf.n + sh.n;
      const VSeg& sg = f ? sf : (h ? sh : sc_);
      const int r = j - (f ? 0 : (h ? sf.n : sf.n + sh.n));
      id[u] = j < tot ? sg.list[sg.a + r] : -1;
      p1[u] = sg.s1;
      p2[u] = sg.s2;
    }
#pragma unroll
    for (int u = 0; u < S3_VB; u++) {
      if (id[u] < 0) continue;
      acc1 = acc1 + ld3(p1[u], id[u]);
      if (TWO && p2[u]) acc2 = acc2 + ld3(p2[u], id[u]);
    }
  }
#pragma unroll
  for (int o = S3_VG / 2; o > 0; o >>= 1) {
    acc1 = acc1 + mk3(__shfl_xor(acc1.x, o, S3_VG), __shfl_xor(acc1.y, o, S3_VG), __shfl_xor(acc1.z, o, S3_VG));
    if (TWO) acc2 = acc2 + mk3(__shfl_xor(acc2.x, o, S3_VG), __shfl_xor(acc2.y, o, S3_VG), __shfl_xor(acc2.z, o, S3_VG));
  }
  sum1 = acc1;
  sum2 = acc2;
}

// Forces and the Hessian diagonal per vertex (system.py:360-365): the
// target term, then the slots of its faces, hinges and contacts in order.
__global__ void k3_diff_vert(int nV, const double* __restrict__ x, const double* __restrict__ target,
                             const double* __restrict__ vor, Cfg c, const int* __restrict__ vf_off,
                             const int* __restrict__ vf, const double* __restrict__ fs, const int* __restrict__ vh_off,
                             const int* __restrict__ vh, const double* __restrict__ hs, const double* __restrict__ hd,
                             const int* __restrict__ vc_off, const int* __restrict__ vc,
                             const double* __restrict__ cs, const double* __restrict__ cq,
                             double* __restrict__ grad, double* __restrict__ diag) {
  // every lane of a group has the same vertex, so the group returns whole
  const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / S3_VG);
  const int g = threadIdx.x & (S3_VG - 1);
  if (i >= nV) return;
  const double w = vor[i] * c.m2gt_k;
  const int af = vf_off[i], ah = vh_off[i], ac = vc_off[i];
  D3 gs, ds;
  gather3<true>(VSeg{vf, af, vf_off[i + 1] - af, fs, nullptr}, VSeg{vh, ah, vh_off[i + 1] - ah, hs, hd},
                VSeg{vc, ac, vc_off[i + 1] - ac, cs, cq}, g, gs, ds);
  const D3 gr = (ld3(x, i) - ld3(target, i)) * w * -1.0 + gs, dg = mk3(w, w, w) + ds;
  if (g == 0) {
    st3(grad, i, gr);
    st3(diag, i, dg);
  }
}

// --------------------------------------------------------------- H dx
// element parts of H dx (system.py:367-371) into the force slots
__global__ void k3_hdx_elem(int nF, int nH, const int* __restrict__ nCp, const double* __restrict__ x,
                            const double* __restrict__ dx,
                            const int* __restrict__ F, const double* __restrict__ W,
                            const double4* __restrict__ invDm, const int4* __restrict__ hinge,
                            const double* __restrict__ hb, const int4* __restrict__ cidx,
                            const double* __restrict__ cd, const double* __restrict__ cdd,
                            const double2* __restrict__ cg, Cfg c, double* __restrict__ fs, double* __restrict__ hs,
                            double* __restrict__ cs) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nF) {
    const int f = (int)i;
    const int a = F[3 * f], b = F[3 * f + 1], e = F[3 * f + 2];
    const double4 m = invDm[f];
    D3 o[3];
    elastic_hdx(ld3(x, a), ld3(x, b), ld3(x, e), ld3(dx, a), ld3(dx, b), ld3(dx, e), W[f], m2(m.x, m.y, m.z, m.w),
                c.mu, c.la, o);
    for (int k = 0; k < 3; k++) st3(fs, 3LL * f + k, o[k]);
  } else if (i < (long long)nF + 12LL * nH) {
    // one lane per row of a hinge's 12x12 (12 consecutive lanes per hinge)
    const long long rr = i - nF;
    const int h = (int)(rr / 12), row = (int)(rr % 12);
    const int4 k = hinge[h];
    const double* Hr = hb + 144LL * h + 12 * row;
    const int id[4] = {k.x, k.y, k.z, k.w};
    double s = 0.0;
    for (int j = 0; j < 4; j++) {
      const D3 d = ld3(dx, id[j]);
      s = s + (Hr[3 * j] * d.x + Hr[3 * j + 1] * d.y + Hr[3 * j + 2] * d.z);
    }
    hs[12LL * h + row] = s;
  } else if (i < (long long)nF + 12LL * nH + *nCp) {
    const int j = (int)(i - nF - 12LL * nH);
    const int4 k = cidx[j];
    if (!(cd[j] < c.d_hat)) {
      for (int r = 0; r < 4; r++) st3(cs, 4LL * j + r, mk3(0.0, 0.0, 0.0));
      return;
    }
    const int id[4] = {k.x, k.y, k.z, k.w};
    D3 dd[4];
    double s = 0.0;
    for (int r = 0; r < 4; r++) {
      dd[r] = ld3(cdd, 4LL * j + r);
      s += dot(dd[r], ld3(dx, id[r]));
    }
    const double g2 = cg[j].y;
    for (int r = 0; r < 4; r++) st3(cs, 4LL * j + r, dd[r] * (s * g2));
  }
}
// out = dx * w + the vertex's face, hinge and contact slots; prod = dx . out
__global__ void k3_hdx_vert(int nV, const double* __restrict__ dx, const double* __restrict__ vor, Cfg c,
                            const int* __restrict__ vf_off, const int* __restrict__ vf, const double* __restrict__ fs,
                            const int* __restrict__ vh_off, const int* __restrict__ vh, const double* __restrict__ hs,
                            const int* __restrict__ vc_off, const int* __restrict__ vc, const double* __restrict__ cs,
                            double* __restrict__ out, double* __restrict__ prod) {
  const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / S3_VG);
  const int g = threadIdx.x & (S3_VG - 1);
  if (i >= nV) return;
  const D3 d = ld3(dx, i);
  const int af = vf_off[i], ah = vh_off[i], ac = vc_off[i];
  D3 os, unused;
  gather3<false>(VSeg{vf, af, vf_off[i + 1] - af, fs, nullptr}, VSeg{vh, ah, vh_off[i + 1] - ah, hs, nullptr},
                 VSeg{vc, ac, vc_off[i + 1] - ac, cs, nullptr}, g, os, unused);
  const D3 o = d * (vor[i] * c.m2gt_k) + os;
  if (g == 0) {
    st3(out, i, o);
    if (prod) prod[i] = dot(d, o);
  }
}

// ----------------------------------------------------------------- CG
// Jacobi-preconditioned CG on H p = grad (cg_solver.py:100-165,
// solver_kernels.py:4-62); z . r of iteration it in SC_ZR0 + (it & 1).
__global__ void k3_cg_init(int nV, const double* __restrict__ grad, const double* __restrict__ diag,
                           double* __restrict__ p, double* __restrict__ r, double* __restrict__ z,
                           double* __restrict__ v, double* __restrict__ prod) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const D3 rr = ld3(grad, i), dg = ld3(diag, i);
  const D3 zz = mk3(rr.x / (dg.x + 1e-6), rr.y / (dg.y + 1e-6), rr.z / (dg.z + 1e-6));
  st3(p, i, mk3(0.0, 0.0, 0.0));
  st3(r, i, rr);
  st3(z, i, zz);
  st3(v, i, zz);
  prod[i] = dot(zz, rr);
}
__global__ void k3_cg_update(int nV, int it, const double* __restrict__ sc, const double* __restrict__ v,
                             const double* __restrict__ Av, const double* __restrict__ diag, double* __restrict__ p,
                             double* __restrict__ r, double* __restrict__ z, double* __restrict__ prod) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const double vAv = sc[SC_VAV], zr_last = sc[SC_ZR0 + (it & 1)];
  D3 rr = ld3(r, i);
  if (vAv > 1e-16) {
    const double alpha = zr_last / vAv;
    st3(p, i, ld3(p, i) + alpha * ld3(v, i));
    rr = rr - alpha * ld3(Av, i);
    st3(r, i, rr);
  }
  const D3 dg = ld3(diag, i);
  const D3 zz = mk3(rr.x / dg.x, rr.y / dg.y, rr.z / dg.z);
  st3(z, i, zz);
  prod[i] = dot(zz, rr);
}
__global__ void k3_cg_next(int nV, int it, const double* __restrict__ sc, const double* __restrict__ z,
                           double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const double s = sc[SC_ZR0 + ((it + 1) & 1)] / sc[SC_ZR0 + (it & 1)];
  st3(v, i, ld3(z, i) + s * ld3(v, i));
}

// One CG iteration after H v (small meshes): v.Av from prod, alpha, the
// updates of p, r and z, z.r, then v, in one workgroup; the same sums
// (block_sum) and updates as k3_sum, k3_cg_update, k3_sum and k3_cg_next,
// so the results are identical.
__global__ void __launch_bounds__(S3_RED) k3_cg_step(int nV, int it, double* __restrict__ sc,
                                                     const double* __restrict__ Av, const double* __restrict__ diag,
                                                     double* __restrict__ v, double* __restrict__ p,
                                                     double* __restrict__ r, double* __restrict__ z,
                                                     double* __restrict__ prod) {
  __shared__ double lds[S3_RED];
  const double vAv = block_sum(prod, nV, lds);
  const double zr_last = sc[SC_ZR0 + (it & 1)];
  for (int i = threadIdx.x; i < nV; i += S3_RED) {
    D3 rr = ld3(r, i);
    if (vAv > 1e-16) {
      const double alpha = zr_last / vAv;
      st3(p, i, ld3(p, i) + alpha * ld3(v, i));
      rr = rr - alpha * ld3(Av, i);
      st3(r, i, rr);
    }
    const D3 dg = ld3(diag, i);
    const D3 zz = mk3(rr.x / dg.x, rr.y / dg.y, rr.z / dg.z);
    st3(z, i, zz);
    prod[i] = dot(zz, rr);
  }
  __syncthreads();
  const double zr_new = block_sum(prod, nV, lds);
  const double sfac = zr_new / zr_last;
  for (int i = threadIdx.x; i < nV; i += S3_RED) st3(v, i, ld3(z, i) + sfac * ld3(v, i));
  if (threadIdx.x == 0) {
    sc[SC_VAV] = vAv;
    sc[SC_ZR0 + ((it + 1) & 1)] = zr_new;
  }
}

// k3_cg_step for meshes of at most 2 S3_RED vertices, each lane's vertices
// held in registers between the phases (one load of the vectors, no reload
// of the products or of z): the same values and sums as k3_cg_step.
constexpr int S3_CG_REG = 2;
__device__ __forceinline__ double block_sum_partial(double s, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = s;
  __syncthreads();
  double t = lds[lane & (S3_RED / 64 - 1)];
#pragma unroll
  for (int o = S3_RED / 128; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  __syncthreads();
  return t;
}
__global__ void __launch_bounds__(S3_RED) k3_cg_step_reg(int nV, int it, double* __restrict__ sc,
                                                         const double* __restrict__ Av,
                                                         const double* __restrict__ diag, double* __restrict__ v,
                                                         double* __restrict__ p, double* __restrict__ r,
                                                         double* __restrict__ z, const double* __restrict__ prod) {
  __shared__ double lds[S3_RED];
  double s0 = 0.0;
  D3 rr[S3_CG_REG], vv[S3_CG_REG], pp[S3_CG_REG], aa[S3_CG_REG], dg[S3_CG_REG];
#pragma unroll
  for (int k = 0; k < S3_CG_REG; k++) {
    const int i = threadIdx.x + k * S3_RED;
    if (i < nV) {
      s0 += prod[i];
      rr[k] = ld3(r, i);
      vv[k] = ld3(v, i);
      pp[k] = ld3(p, i);
      aa[k] = ld3(Av, i);
      dg[k] = ld3(diag, i);
    }
  }
  const double vAv = block_sum_partial(s0, lds);
  const double zr_last = sc[SC_ZR0 + (it & 1)];
  double s1 = 0.0;
  D3 zz[S3_CG_REG];
#pragma unroll
  for (int k = 0; k < S3_CG_REG; k++) {
    const int i = threadIdx.x + k * S3_RED;
    if (i < nV) {
      if (vAv > 1e-16) {
        const double alpha = zr_last / vAv;
        pp[k] = pp[k] + alpha * vv[k];
        rr[k] = rr[k] - alpha * aa[k];
        st3(p, i, pp[k]);
        st3(r, i, rr[k]);
      }
      zz[k] = mk3(rr[k].x / dg[k].x, rr[k].y / dg[k].y, rr[k].z / dg[k].z);
      st3(z, i, zz[k]);
      s1 += dot(zz[k], rr[k]);
    }
  }
  const double zr_new = block_sum_partial(s1, lds);
  const double sfac = zr_new / zr_last;
#pragma unroll
  for (int k = 0; k < S3_CG_REG; k++) {
    const int i = threadIdx.x + k * S3_RED;
    if (i < nV) st3(v, i, zz[k] + sfac * vv[k]);
  }
  if (threadIdx.x == 0) {
    sc[SC_VAV] = vAv;
    sc[SC_ZR0 + ((it + 1) & 1)] = zr_new;
  }
}

// ------------------------------------------------- clamp, CCD, search
// the step limited to the detection radius around q_det
// (solver_kernels.py:76-94)
__global__ void k3_clamp(int nV, const double* __restrict__ qn, const double* __restrict__ qdet, double radius,
                         double* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const D3 dq = ld3(qn, i) + ld3(p, i) - ld3(qdet, i);
  if (len(dq) > radius) st3(p, i, nrm(dq) * radius + ld3(qdet, i) - ld3(qn, i));
}
__global__ void k3_ccd(const int* __restrict__ nCp, const double* __restrict__ q, const double* __restrict__ p,
                       const int4* __restrict__ cidx, const int2* __restrict__ ctyp, Cfg c, double* __restrict__ sc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *nCp) return;
  const int4 k = cidx[j];
  const int id[4] = {k.x, k.y, k.z, k.w};
  D3 x[4], v[4];
  for (int r = 0; r < 4; r++) {
    x[r] = ld3(q, id[r]);
    v[r] = ld3(p, id[r]);
  }
  const double t = accd_pair(x, v, ctyp[j].x, c.slack, c.thick, c.ccd_iters, c.ee_thres);
  atomicMin((unsigned long long*)(sc + SC_CCD), (unsigned long long)__double_as_longlong(t));
}
// Line search step n (system.py:395-405, solver_kernels.py:64-74): until
// the energy has dropped below the energy before the search, the positions
// take the step ccd_step * 2^-n; step 0 also saves that energy.
__global__ void k3_ls_pos(int nV, int n, double* __restrict__ sc, const double* __restrict__ qn,
                          const double* __restrict__ p, double* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  if (n == 0 && i == 0) sc[SC_EPREV] = sc[SC_E];
  if (n > 0 && sc[SC_E] < sc[SC_EPREV]) return;
  const double ccd = __longlong_as_double((long long)((const unsigned long long*)sc)[SC_CCD]);
  const double a = ccd * ldexp(1.0, -n);
  st3(q, i, ld3(qn, i) + a * ld3(p, i));
  if (i == 0) sc[SC_LS] = (double)n;
}
__global__ void k3_set_u64(unsigned long long* a, unsigned long long v) { *a = v; }
__global__ void k3_set_i32(int* a, int v) { *a = v; }

}  // namespace s3
