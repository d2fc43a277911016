// CPU restatement, in float64, of the reference's stage 3 (safe projection):
// the Newton / CG / CCD loop of
//   simp_cuda/safe_project/src/pamo_safe_project/system.py:17-470,
//   cg_solver.py:31-165, kernels/solver_kernels.py:1-94,
// with the four energies of the default configuration (config.py:26-35):
//   Mesh2GTDistanceBvhEnergyCalculator  energy.py:193-211, distance_energy.py:6-47,
//                                       geometry_kernels.py:91-104
//   ElasticEnergyCalculator             energy.py:337-432, elastic_energy.py:1-220
//   HingeEnergyCalculator               energy.py:589-700, hinge_energy.py:1-274,
//                                       utils_kernels.py:20-41
//   CollisionBvhEnergyCalculator        energy.py:705-946, collision_energy.py:1-409,
//                                       contact_detection.py:120-216, ccd_kernels.py:8-110
// and the distance classification of distance_kernels/distance_kernels.py:1-545.
//
// TEST INFRASTRUCTURE ONLY.  Loaded by tests/ and bench.py's CPU leg through
// oracle/stage3.py, always as the checker; the product package never loads it.
//
// How each part is pinned (tests/test_stage3_oracle.py):
//   distances and their types   the closest point on a triangle by Ericson's
//                               region test and a dense (s, t) search for
//                               segment pairs; special cases by hand
//   distance gradients          central differences of the distance
//   barrier b, b', b''          closed form at d_hat and central differences
//   elastic energy / force / Hv a pure stretch (closed form), differences of
//                               the energy and of the force (in-plane dx)
//   hinge angle / force / H     flaps of known dihedral angle, differences of
//                               the energy and of the force
//   SPD projection              numpy's symmetric eigensolver on the same 9x9
//   Voronoi areas               equilateral, right and obtuse triangles by hand,
//                               partition of the total area
//   ACCD step                   a point moving straight at a plane: the step
//                               stops short of the contact, by the slackness
//   CG                          numpy.linalg.solve on a small SPD system
//   Newton step                 energy does not increase; the line search
//                               takes the first halving that lowers it
// Parity unpinned: the 9x9 SPD projection of the hinge blocks calls
// wp.spd_project_blocks of the authors' Warp fork (a git submodule absent
// from the reference); it is restated as it_max cyclic Jacobi sweeps with the
// negative eigenvalues clamped to zero (definition S3-D1).
//
// Definitions where the reference's order is not fixed (float atomics):
// every sum here is taken in float64 in element order; contacts are listed
// point-triangle pairs first, by (vertex, triangle), then edge-edge pairs by
// (edge, edge), edges numbered in ascending (min, max) vertex order.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace s3o {

struct V {
  double x, y, z;
};
static inline V mk(double x, double y, double z) { return V{x, y, z}; }
static inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
static inline V operator*(double s, V a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline V operator*(V a, double s) { return mk(s * a.x, s * a.y, s * a.z); }
static inline V operator/(V a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double len2(V a) { return dot(a, a); }
static inline double len(V a) { return std::sqrt(dot(a, a)); }
// wp.normalize: the zero vector stays zero
static inline V nrm(V a) {
  double l = len(a);
  return l > 0.0 ? a / l : mk(0, 0, 0);
}
static inline V cmul(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }

// ---------------------------------------------------------------- types
enum { PT = 0, PP01 = 1, PP02 = 2, PE012 = 3, PP03 = 4, PE031 = 5, PE023 = 6 };   // defs.py:12-19
enum { EPP02 = 0, EPP03 = 1, EPP12 = 2, EPP13 = 3, EPE023 = 4, EPE123 = 5, EPE201 = 6, EPE301 = 7, EEE = 8 };   // :22-31
enum { BT_PT = 3, BT_EE = 4 };   // defs.py:1-9

// distance_kernels.py:8-27
static double d_pt_plane(V x0, V x1, V x2, V x3) { return std::fabs(dot(x0 - x1, nrm(cross(x2 - x1, x3 - x1)))); }
static double d_ee_line(V x0, V x1, V x2, V x3) { return std::fabs(dot(x0 - x2, nrm(cross(x1 - x0, x3 - x2)))); }
static double d_pe(V x0, V x1, V x2) { return len(cross(x1 - x0, x2 - x0)) / len(x1 - x2); }
static double d_pp(V x0, V x1) { return len(x1 - x0); }

// distance_kernels.py:79-205
static int pt_classify(V x0, V x1, V x2, V x3) {
  V x10 = x0 - x1, x12 = x2 - x1, x13 = x3 - x1;
  double area_sq = len2(cross(x12, x13));
  if (area_sq <= 1e-6 * (len2(x12) * len2(x13)) || area_sq <= 1e-16) {
    double d1 = len(x10), d2 = len(x0 - x2), d3 = len(x0 - x3);
    if (d1 < d2) return d1 < d3 ? PP01 : PP03;
    return d2 < d3 ? PP02 : PP03;
  }
  V n = nrm(cross(x12, x13));
  V xp = x0 - dot(n, x10) * n;
  x10 = xp - x1;
  V x20 = xp - x2, x30 = xp - x3;
  double a00 = dot(x12, x12), a01 = dot(x12, x13), a11 = dot(x13, x13);
  double b0 = dot(x10, x12), b1 = dot(x10, x13);
  double det = a00 * a11 - a01 * a01;
  double beta = b0 * a11 - b1 * a01, gamma = a00 * b1 - b0 * a01, alpha = det - beta - gamma;
  int nneg = (alpha < 0.0) + (beta < 0.0) + (gamma < 0.0);
  if (nneg == 0) return PT;
  if (nneg == 1) {
    if (alpha < 0.0) {   // edge x2x3, with the unprojected point
      V x23 = x3 - x2;
      double r = dot(x23, x0 - x2);
      if (r < 0.0) return PP02;
      if (r > dot(x23, x23)) return PP03;
      return PE023;
    }
    if (beta < 0.0) {   // edge x3x1
      V x31 = x1 - x3;
      double r = dot(x31, x0 - x3);
      if (r < 0.0) return PP03;
      if (r > dot(x31, x31)) return PP01;
      return PE031;
    }
    double r = dot(x12, x10);   // edge x1x2, projected point
    if (r < 0.0) return PP01;
    if (r > dot(x12, x12)) return PP02;
    return PE012;
  }
  if (alpha >= 0.0) {
    double d12 = dot(x12, x10), d13 = dot(x13, x10);
    if (d12 > 0.0 && d13 <= 0.0) return dot(x12, x20) > 0.0 ? PP02 : PE012;
    if (d13 > 0.0 && d12 <= 0.0) return dot(x13, x30) > 0.0 ? PP03 : PE031;
    return PP01;
  }
  if (beta >= 0.0) {
    V x23 = x3 - x2, x21 = x1 - x2;
    double d23 = dot(x23, x20), d21 = dot(x21, x20);
    if (d23 > 0.0 && d21 <= 0.0) return dot(x23, x30) > 0.0 ? PP03 : PE023;
    if (d21 > 0.0 && d23 <= 0.0) return dot(x21, x10) > 0.0 ? PP01 : PE012;
    return PP02;
  }
  V x31 = x1 - x3, x32 = x2 - x3;
  double d31 = dot(x31, x30), d32 = dot(x32, x30);
  if (d31 > 0.0 && d32 <= 0.0) return dot(x31, x10) > 0.0 ? PP01 : PE031;
  if (d32 > 0.0 && d31 <= 0.0) return dot(x32, x20) > 0.0 ? PP02 : PE023;
  return PP03;
}

// distance_kernels.py:228-245
static double pt_distance(V x0, V x1, V x2, V x3, int t) {
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

// Gradients of the distances above, from the envelope theorem: the distance
// to an affine set is |x0 - y*| at the foot y*; moving a defining point moves
// the set with weight equal to that point's affine coordinate of y*.
static void g_pp(V a, V b, int ia, int ib, V g[4]) {
  g[ia] = nrm(a - b);
  g[ib] = nrm(b - a);
}
static void g_pe(V p, V a, V b, int ip, int ia, int ib, V g[4]) {
  V ab = b - a;
  double t = dot(p - a, ab) / dot(ab, ab);
  V u = nrm(p - (a + t * ab));
  g[ip] = u;
  g[ia] = -(1.0 - t) * u;
  g[ib] = -t * u;
}
static void g_pt_plane(V x0, V x1, V x2, V x3, V g[4]) {
  V x12 = x2 - x1, x13 = x3 - x1, x10 = x0 - x1;
  V n = nrm(cross(x12, x13));
  double s = dot(x10, n);
  V u = s > 0.0 ? n : (s < 0.0 ? -n : mk(0, 0, 0));
  double a00 = dot(x12, x12), a01 = dot(x12, x13), a11 = dot(x13, x13);
  double b0 = dot(x10, x12), b1 = dot(x10, x13);
  double det = a00 * a11 - a01 * a01;
  double beta = (b0 * a11 - b1 * a01) / det, gamma = (a00 * b1 - b0 * a01) / det;
  double alpha = 1.0 - beta - gamma;
  g[0] = u;
  g[1] = -alpha * u;
  g[2] = -beta * u;
  g[3] = -gamma * u;
}
static void g_ee_line(V x0, V x1, V x2, V x3, V g[4]) {
  V u = x1 - x0, v = x3 - x2, w = x0 - x2;
  V n = nrm(cross(u, v));
  double s = dot(w, n);
  V un = s > 0.0 ? n : (s < 0.0 ? -n : mk(0, 0, 0));
  double A = dot(u, u), B = dot(u, v), C = dot(v, v), D = dot(u, w), E = dot(v, w);
  double den = A * C - B * B;
  double a = (B * E - C * D) / den, b = (A * E - B * D) / den;
  g[0] = (1.0 - a) * un;
  g[1] = a * un;
  g[2] = -(1.0 - b) * un;
  g[3] = -b * un;
}
static void pt_grad(V x0, V x1, V x2, V x3, int t, V g[4]) {
  for (int k = 0; k < 4; k++) g[k] = mk(0, 0, 0);
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

// distance_kernels.py:309-430
static int ee_classify(V x0, V x1, V x2, V x3, double sin_thres) {
  V e0 = x0 - x1, e1 = x2 - x3;
  double e0s = dot(e0, e0), e1s = dot(e1, e1);
  double a00 = e0s, a01 = -dot(e0, e1), a10 = dot(e0, e1), a11 = -e1s;
  double b0 = dot(x3 - x1, e0), b1 = dot(x3 - x1, e1);
  double mu0 = -(b0 * a11 - b1 * a01), mu1 = -(a00 * b1 - a10 * b0), detA = -(a00 * a11 - a01 * a10);
  double esin = len(cross(nrm(e0), nrm(e1)));
  if (mu0 >= 0 && mu0 <= detA && mu1 >= 0 && mu1 <= detA && esin >= sin_thres && e0s > 1e-15 && e1s > 1e-15)
    return EEE;
  double best = 1e9;
  int type = -1;
  auto take = [&](double d, int t) {
    if (d < best) { best = d; type = t; }
  };
  double t0 = dot(x0 - x3, e1);
  if (t0 >= e1s) take(d_pp(x0, x2), EPP02);
  else if (t0 <= 0) take(d_pp(x0, x3), EPP03);
  else take(d_pe(x0, x2, x3), EPE023);
  double t1 = dot(x1 - x3, e1);
  if (t1 >= e1s) take(d_pp(x1, x2), EPP12);
  else if (t1 <= 0) take(d_pp(x1, x3), EPP13);
  else take(d_pe(x1, x2, x3), EPE123);
  double t2 = dot(x2 - x1, e0);
  if (t2 >= e0s) take(d_pp(x2, x0), EPP02);
  else if (t2 <= 0) take(d_pp(x2, x1), EPP12);
  else take(d_pe(x2, x0, x1), EPE201);
  double t3 = dot(x3 - x1, e0);
  if (t3 >= e0s) take(d_pp(x3, x0), EPP03);
  else if (t3 <= 0) take(d_pp(x3, x1), EPP13);
  else take(d_pe(x3, x0, x1), EPE301);
  return type;
}

// distance_kernels.py:433-452
static double ee_distance(V x0, V x1, V x2, V x3, int t) {
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
static void ee_grad(V x0, V x1, V x2, V x3, int t, V g[4]) {
  for (int k = 0; k < 4; k++) g[k] = mk(0, 0, 0);
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

// Closest point of a triangle to p (Ericson, Real-Time Collision Detection,
// 5.1.5), the query of wp.mesh_query_point_no_sign + mesh_eval_position
// (geometry_kernels.py:91-104).
static V closest_on_tri(V p, V a, V b, V c) {
  V ab = b - a, ac = c - a, ap = p - a;
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  V bp = p - b;
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + (d1 / (d1 - d3)) * ab;
  V cp = p - c;
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + (d2 / (d2 - d6)) * ac;
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
  double den = 1.0 / (va + vb + vc);
  double v = vb * den, w = vc * den;
  return a + v * ab + w * ac;
}

// collision_energy.py:8-27 (wp.select(c, a, b) is c ? b : a)
static double bar(double d, double dh) { return d >= dh ? 0.0 : -(d - dh) * (d - dh) * std::log(d / dh); }
static double bar1(double d, double dh) {
  double t = d - dh;
  return d >= dh ? 0.0 : -2.0 * t * std::log(d / dh) - t * t / d;
}
static double bar2(double d, double dh) {
  return d >= dh ? 0.0 : -2.0 * std::log(d / dh) - 2.0 * (d - dh) / d - 1.0 + (dh * dh) / (d * d);
}

// ------------------------------------------------------------- elastic
// 2x2 row-major: m[0]=(0,0) m[1]=(0,1) m[2]=(1,0) m[3]=(1,1)
struct M2 {
  double a, b, c, d;
};
static inline M2 mul(M2 x, M2 y) {
  return M2{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
static inline M2 tr(M2 x) { return M2{x.a, x.c, x.b, x.d}; }
static inline M2 add(M2 x, M2 y) { return M2{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
static inline M2 scl(double s, M2 x) { return M2{s * x.a, s * x.b, s * x.c, s * x.d}; }
static inline double trace(M2 x) { return x.a + x.d; }
static inline M2 inv(M2 x) {   // wp.inverse: zero on a zero determinant
  double det = x.a * x.d - x.b * x.c;
  if (det == 0.0) return M2{0, 0, 0, 0};
  return M2{x.d / det, -x.b / det, -x.c / det, x.a / det};
}
static void frame(V x0, V x1, V x2, V& b1, V& b2) {
  V x01 = x1 - x0, x02 = x2 - x0;
  b1 = nrm(x01);
  b2 = nrm(x02 - dot(x02, b1) * b1);
}
static M2 local(V b1, V b2, V x01, V x02) { return M2{dot(b1, x01), dot(b1, x02), dot(b2, x01), dot(b2, x02)}; }
// elastic_energy.py:7-37
static void elastic_rest(V x0, V x1, V x2, double& W, M2& invDm) {
  V x01 = x1 - x0, x02 = x2 - x0;
  W = 0.5 * len(cross(x01, x02));
  V b1, b2;
  frame(x0, x1, x2, b1, b2);
  invDm = inv(local(b1, b2, x01, x02));
}
// elastic_energy.py:40-83; 86-145 (force f = coeff * dpsi/dx); 148-220 (H dx)
static double elastic_energy(V x0, V x1, V x2, double W, M2 invDm, double mu, double la) {
  V b1, b2;
  frame(x0, x1, x2, b1, b2);
  M2 F = mul(local(b1, b2, x1 - x0, x2 - x0), invDm);
  M2 E = scl(0.5, add(mul(tr(F), F), M2{-1, 0, 0, -1}));
  double te = trace(E);
  return W * (mu * trace(mul(E, tr(E))) + 0.5 * la * te * te);
}
static void elastic_force(V x0, V x1, V x2, double W, M2 invDm, double mu, double la, double coeff, V f[3]) {
  V b1, b2;
  frame(x0, x1, x2, b1, b2);
  M2 F = mul(local(b1, b2, x1 - x0, x2 - x0), invDm);
  M2 E = scl(0.5, add(mul(tr(F), F), M2{-1, 0, 0, -1}));
  double te = trace(E);
  M2 P = mul(F, add(scl(2.0 * mu, E), M2{la * te, 0, 0, la * te}));
  M2 H = mul(scl(coeff * W, P), tr(invDm));
  f[1] = H.a * b1 + H.c * b2;
  f[2] = H.b * b1 + H.d * b2;
  f[0] = -f[1] - f[2];
}
static void elastic_hdx(V x0, V x1, V x2, V d0, V d1, V d2, double W, M2 invDm, double mu, double la, V f[3]) {
  V b1, b2;
  frame(x0, x1, x2, b1, b2);
  M2 F = mul(local(b1, b2, x1 - x0, x2 - x0), invDm);
  M2 dF = mul(local(b1, b2, d1 - d0, d2 - d0), invDm);
  M2 E = scl(0.5, add(mul(tr(F), F), M2{-1, 0, 0, -1}));
  double te = trace(E);
  M2 dE = scl(0.5, add(mul(tr(dF), F), mul(tr(F), dF)));
  double tde = trace(dE);
  M2 dP = add(mul(dF, add(scl(2.0 * mu, E), M2{la * te, 0, 0, la * te})),
              mul(F, add(scl(2.0 * mu, dE), M2{la * tde, 0, 0, la * tde})));
  M2 dH = mul(scl(W, dP), tr(invDm));
  f[1] = dH.a * b1 + dH.c * b2;
  f[2] = dH.b * b1 + dH.d * b2;
  f[0] = -f[1] - f[2];
}

// --------------------------------------------------------------- hinge
// 3x3 row-major
struct M3 {
  double m[9];
};
static inline M3 outer(V a, V b) {
  return M3{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
static inline M3 operator+(M3 a, M3 b) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = a.m[i] + b.m[i];
  return r;
}
static inline M3 operator-(M3 a, M3 b) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = a.m[i] - b.m[i];
  return r;
}
static inline M3 operator*(double s, M3 a) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = s * a.m[i];
  return r;
}
static inline M3 T3(M3 a) {
  return M3{{a.m[0], a.m[3], a.m[6], a.m[1], a.m[4], a.m[7], a.m[2], a.m[5], a.m[8]}};
}
static inline M3 sym(M3 a) { return a + T3(a); }
static inline V mv(M3 a, V v) {
  return mk(a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z);
}

// hinge_energy.py:4-13: signed dihedral angle of the flaps (x0; x1, x2) and (x3; x1, x2)
static double hinge_angle(V x0, V x1, V x2, V x3) {
  V e12 = nrm(x2 - x1);
  V n1 = nrm(cross(e12, x0 - x1));
  V n2 = nrm(cross(x3 - x1, e12));
  return std::atan2(dot(cross(n1, n2), e12), dot(n1, n2));
}

// hinge_energy.py:89-255 (Tamstorf and Grinspun 2013, "Discrete bending
// forces and their Jacobians"): force (coeff * dE/dx) and the 4x4 blocks of
// the Hessian, blocks[i][j] for vertices i, j of (x0, x1, x2, x3).
static void hinge_diff(V x0, V x1, V x2, V x3, double rest, double elen, double k, double coeff, V g[4],
                       M3 blk[4][4]) {
  V e0 = x2 - x1, e1 = x0 - x2, e2 = x0 - x1, e1_ = x3 - x2, e2_ = x3 - x1;
  double e0s = len2(e0), e1s = len2(e1), e2s = len2(e2), e1s_ = len2(e1_), e2s_ = len2(e2_);
  double l0 = std::sqrt(e0s), l1 = std::sqrt(e1s), l2 = std::sqrt(e2s), l1_ = std::sqrt(e1s_), l2_ = std::sqrt(e2s_);
  V u0 = nrm(e0), u1 = nrm(e1), u2 = nrm(e2), u1_ = nrm(e1_), u2_ = nrm(e2_);
  double ca1 = dot(u2, u0), ca2 = -dot(u1, u0), ca1_ = dot(u2_, u0), ca2_ = -dot(u1_, u0);
  double h0 = len(cross(e1, e2)) / l0, h1 = len(cross(e0, e2)) / l1, h2 = len(cross(e0, e1)) / l2;
  double h0_ = len(cross(e1_, e2_)) / l0, h1_ = len(cross(e0, e2_)) / l1_, h2_ = len(cross(e0, e1_)) / l2_;
  V n = nrm(cross(e0, e2)), n_ = nrm(cross(e2_, e0));
  double th = std::atan2(dot(cross(n, n_), e0) / l0, dot(n, n_));
  double dE = k * (th - rest) * elen, d2E = k * elen;
  V t1 = n * (ca2 / h1) + n_ * (ca2_ / h1_);
  V t2 = n * (ca1 / h2) + n_ * (ca1_ / h2_);
  V t0 = n * (-1.0 / h0), t3 = n_ * (-1.0 / h0_);
  g[0] = dE * t0 * coeff;
  g[1] = dE * t1 * coeff;
  g[2] = dE * t2 * coeff;
  g[3] = dE * t3 * coeff;
  V m0 = nrm(cross(e0, n)), m1 = nrm(cross(e1, n)), m2 = nrm(cross(n, e2));
  V m0_ = nrm(cross(n_, e0)), m1_ = nrm(cross(n_, e1_)), m2_ = nrm(cross(e2_, n_));
  M3 M0 = outer(n, m0), M1 = outer(n, m1), M2_ = outer(n, m2);
  M3 M0_ = outer(n_, m0_), M1_ = outer(n_, m1_), M2__ = outer(n_, m2_);
  M3 N0 = (1.0 / e0s) * M0, N0_ = (1.0 / e0s) * M0_;
  M3 P10 = (ca1 / (h1 * h0)) * T3(M0), P11 = (ca1 / (h1 * h1)) * T3(M1), P12 = (ca1 / (h1 * h2)) * T3(M2_);
  M3 P20 = (ca2 / (h2 * h0)) * T3(M0), P21 = (ca2 / (h2 * h1)) * T3(M1), P22 = (ca2 / (h2 * h2)) * T3(M2_);
  M3 P10_ = (ca1_ / (h1_ * h0_)) * T3(M0_), P11_ = (ca1_ / (h1_ * h1_)) * T3(M1_),
     P12_ = (ca1_ / (h1_ * h2_)) * T3(M2__);
  M3 P20_ = (ca2_ / (h2_ * h0_)) * T3(M0_), P21_ = (ca2_ / (h2_ * h1_)) * T3(M1_),
     P22_ = (ca2_ / (h2_ * h2_)) * T3(M2__);
  M3 Q0 = (1.0 / (h0 * h0)) * M0, Q1 = (1.0 / (h0 * h1)) * M1, Q2 = (1.0 / (h0 * h2)) * M2_;
  M3 Q0_ = (1.0 / (h0_ * h0_)) * M0_, Q1_ = (1.0 / (h0_ * h1_)) * M1_, Q2_ = (1.0 / (h0_ * h2_)) * M2__;
  blk[0][0] = (-dE) * sym(Q0) + d2E * outer(t0, t0);
  blk[3][3] = (-dE) * sym(Q0_) + d2E * outer(t3, t3);
  blk[1][1] = dE * (sym(P11) - N0 + sym(P11_) - N0_) + d2E * outer(t1, t1);
  blk[2][2] = dE * (sym(P22) - N0 + sym(P22_) - N0_) + d2E * outer(t2, t2);
  blk[1][0] = dE * (P10 - Q1) + d2E * outer(t1, t0);
  blk[2][0] = dE * (P20 - Q2) + d2E * outer(t2, t0);
  blk[1][3] = dE * (P10_ - Q1_) + d2E * outer(t1, t3);
  blk[2][3] = dE * (P20_ - Q2_) + d2E * outer(t2, t3);
  blk[1][2] = dE * (P12 + T3(P21) + N0 + P12_ + T3(P21_) + N0_) + d2E * outer(t1, t2);
  blk[0][3] = d2E * outer(t0, t3);
  blk[0][1] = T3(blk[1][0]);
  blk[0][2] = T3(blk[2][0]);
  blk[3][1] = T3(blk[1][3]);
  blk[3][2] = T3(blk[2][3]);
  blk[2][1] = T3(blk[1][2]);
  blk[3][0] = T3(blk[0][3]);
}

// S3-D1: the 9x9 matrix of blocks [0..2][0..2] projected on the PSD cone by
// it_max cyclic Jacobi sweeps (Golub and Van Loan, 8.5.2) with negative
// eigenvalues set to zero; row and column 3 rebuilt from zero row sums
// (utils_kernels.py:20-41).
static void spd_project(M3 blk[4][4], int it_max) {
  double A[9][9], Q[9][9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) A[3 * i + a][3 * j + b] = blk[i][j].m[3 * a + b];
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) Q[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < it_max; sweep++)
    for (int p = 0; p < 8; p++)
      for (int q = p + 1; q < 9; q++) {
        double apq = A[p][q];
        if (apq == 0.0) continue;
        double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
        double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau)) : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < 9; k++) {   // A <- A J
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 9; k++) {   // A <- J^T A
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 9; k++) {   // Q <- Q J
          double qkp = Q[k][p], qkq = Q[k][q];
          Q[k][p] = c * qkp - s * qkq;
          Q[k][q] = s * qkp + c * qkq;
        }
      }
  double lam[9];
  for (int i = 0; i < 9; i++) lam[i] = A[i][i] > 0.0 ? A[i][i] : 0.0;
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) {
      double s = 0.0;
      for (int k = 0; k < 9; k++) s += Q[i][k] * lam[k] * Q[j][k];
      blk[i / 3][j / 3].m[3 * (i % 3) + (j % 3)] = s;
    }
  for (int j = 0; j < 3; j++) blk[3][j] = (-1.0) * (blk[0][j] + blk[1][j] + blk[2][j]);
  for (int j = 0; j < 3; j++) blk[j][3] = T3(blk[3][j]);
  blk[3][3] = (-1.0) * (blk[0][3] + blk[1][3] + blk[2][3]);
}

// ---------------------------------------------------------------- ACCD
// ccd_kernels.py:8-110 for one contact (x: positions, v: search direction);
// returns the step it limits the search to, or 1 where it does not.
static double accd_pair(V x[4], V v[4], int type1, double slack, double thick, int max_iters, double ee_thres) {
  V va = (v[0] + v[1] + v[2] + v[3]) / 4.0;
  V w[4];
  for (int k = 0; k < 4; k++) w[k] = v[k] - va;
  double lv = type1 == BT_PT ? len(w[0]) + std::max(len(w[1]), std::max(len(w[2]), len(w[3])))
                             : std::max(len(w[0]), len(w[1])) + std::max(len(w[2]), len(w[3]));
  if (lv < 1e-12) return 1.0;
  V p[4] = {x[0], x[1], x[2], x[3]};
  auto dist = [&]() {
    if (type1 == BT_PT) return pt_distance(p[0], p[1], p[2], p[3], pt_classify(p[0], p[1], p[2], p[3]));
    return ee_distance(p[0], p[1], p[2], p[3], ee_classify(p[0], p[1], p[2], p[3], ee_thres));
  };
  auto gap = [&](double d) {
    double ds = d * d;
    double r = (ds - thick * thick) / (std::sqrt(ds) + thick);
    return r > 0.0 ? r : 0.0;
  };
  double dmt = gap(dist());
  double g = (1.0 - slack) * dmt;
  double t = 0.0, tl = slack * dmt / lv;
  for (int i = 0;; i++) {
    for (int k = 0; k < 4; k++) p[k] = p[k] + tl * w[k];
    dmt = gap(dist());
    if ((i > 0 && dmt <= g) || i >= max_iters) return t;
    t = t + tl;
    if (t > 1.0) return 1.0;
    tl = 0.8 * dmt / lv;
  }
}

// ----------------------------------------------------- mesh attributes
// Mixed Voronoi area per vertex (Meyer, Desbrun, Schroeder and Barr 2003),
// the diagonal of igl::massmatrix(V, F, MASSMATRIX_TYPE_VORONOI)
// (system.py:293-295): a non-obtuse triangle gives vertex i
// (|e_ij|^2 cot(angle k) + |e_ik|^2 cot(angle j)) / 8; an obtuse one gives
// half its area to the obtuse corner and a quarter to the others.
static void voronoi(const std::vector<V>& X, const std::vector<std::array<int, 3>>& F, std::vector<double>& out) {
  out.assign(X.size(), 0.0);
  for (auto& f : F) {
    V p[3] = {X[f[0]], X[f[1]], X[f[2]]};
    double A = 0.5 * len(cross(p[1] - p[0], p[2] - p[0]));
    double c[3], cot[3];
    for (int i = 0; i < 3; i++) {
      V a = p[(i + 1) % 3] - p[i], b = p[(i + 2) % 3] - p[i];
      c[i] = dot(a, b);
      cot[i] = c[i] / len(cross(a, b));
    }
    if (c[0] < 0.0 || c[1] < 0.0 || c[2] < 0.0) {
      for (int i = 0; i < 3; i++) out[f[i]] += c[i] < 0.0 ? 0.5 * A : 0.25 * A;
      continue;
    }
    for (int i = 0; i < 3; i++) {
      int j = (i + 1) % 3, k = (i + 2) % 3;
      out[f[i]] += (len2(p[j] - p[i]) * cot[k] + len2(p[k] - p[i]) * cot[j]) / 8.0;
    }
  }
}
// smallest corner angle in degrees (igl.internal_angles, system.py:288-292)
static double min_angle_deg(const std::vector<V>& X, const std::vector<std::array<int, 3>>& F) {
  double m = 180.0;
  for (auto& f : F)
    for (int i = 0; i < 3; i++) {
      V a = X[f[(i + 1) % 3]] - X[f[i]], b = X[f[(i + 2) % 3]] - X[f[i]];
      double c = dot(a, b) / (len(a) * len(b));
      if (c != c) continue;
      c = std::max(-1.0, std::min(1.0, c));
      m = std::min(m, std::acos(c) * 180.0 / M_PI);
    }
  return m;
}

// --------------------------------------------------------------- system
struct Cfg {
  int n_newton = 10, n_cg = 40, n_ls = 10, spd_iters = 3, ccd_iters = 100;
  double d_hat = 1e-3, kappa = 1e2, radius = 1e-2, ee_thres = 1e-3, slack = 0.7, thick = 1e-6;
  double young = 1e-1, poisson = 0.0, hinge_k = 1e-2, m2gt_k = 1e3, min_angle = 0.0;
  long long max_blocks = 1LL << 25;
};

struct Contact {
  int type1, i[4], type2;
  double d;
  V dd[4];
};

struct System {
  Cfg c;
  double mu = 0, la = 0, radius = 0;
  int nV = 0;
  std::vector<V> q, q_rest, q_newton, q_det, target, grad, diag, p, gtV;
  std::vector<std::array<int, 3>> F, gtF;
  std::vector<std::array<int, 2>> E;
  std::vector<double> vor, W;
  std::vector<M2> invDm;
  std::vector<std::array<int, 4>> hinge;
  std::vector<double> rest_angle, rest_len;
  std::vector<std::array<std::array<M3, 4>, 4>> hblk;
  std::vector<Contact> con;
  double energy = 0, energy_prev = 0, ccd_step = 1.0;
  long long ls_taken = 0;
  bool do_nothing = false;

  void setup(const Cfg& cfg) {
    c = cfg;
    mu = c.young / (2.0 * (1.0 + c.poisson));
    la = c.young * c.poisson / ((1.0 + c.poisson) * (1.0 - 2.0 * c.poisson));
    radius = c.radius;
  }

  // system.py:210-320
  int register_mesh(const double* gv, int ngv, const int* gf, int ngf, const double* v, int nv, const int* f, int nf) {
    nV = nv;
    q.resize(nv);
    for (int i = 0; i < nv; i++) q[i] = mk(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    q_rest = q_newton = q_det = target = q;
    grad.assign(nv, mk(0, 0, 0));
    diag = p = grad;
    F.resize(nf);
    for (int i = 0; i < nf; i++) F[i] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
    gtV.resize(ngv);
    for (int i = 0; i < ngv; i++) gtV[i] = mk(gv[3 * i], gv[3 * i + 1], gv[3 * i + 2]);
    gtF.resize(ngf);
    for (int i = 0; i < ngf; i++) gtF[i] = {gf[3 * i], gf[3 * i + 1], gf[3 * i + 2]};
    // unique edges, ascending
    E.clear();
    for (auto& t : F)
      for (int k = 0; k < 3; k++) {
        int a = t[k], b = t[(k + 1) % 3];
        E.push_back({std::min(a, b), std::max(a, b)});
      }
    std::sort(E.begin(), E.end());
    E.erase(std::unique(E.begin(), E.end()), E.end());
    voronoi(q, F, vor);
    do_nothing = min_angle_deg(q, F) <= c.min_angle;
    W.resize(nf);
    invDm.resize(nf);
    for (int i = 0; i < nf; i++) elastic_rest(q[F[i][0]], q[F[i][1]], q[F[i][2]], W[i], invDm[i]);
    // hinges (hinge_energy.py:16-50): face t0 holds the directed edge a->b
    // with a < b, face t1 holds b->a; ordered by (a, b)
    std::vector<std::array<int, 4>> half;   // (a, b, face, corner of the opposite vertex)
    for (int i = 0; i < nf; i++)
      for (int k = 0; k < 3; k++) half.push_back({F[i][k], F[i][(k + 1) % 3], i, (k + 2) % 3});
    std::vector<std::array<int, 4>> fwd, bwd;
    for (auto& h : half) (h[0] < h[1] ? fwd : bwd).push_back(h);
    std::sort(fwd.begin(), fwd.end());
    std::sort(bwd.begin(), bwd.end(), [](const std::array<int, 4>& x, const std::array<int, 4>& y) {
      return std::make_pair(std::make_pair(x[1], x[0]), x[2]) < std::make_pair(std::make_pair(y[1], y[0]), y[2]);
    });
    hinge.clear();
    rest_angle.clear();
    rest_len.clear();
    size_t j = 0;
    for (auto& h : fwd) {
      while (j < bwd.size() && std::make_pair(bwd[j][1], bwd[j][0]) < std::make_pair(h[0], h[1])) j++;
      for (size_t k = j; k < bwd.size() && bwd[k][1] == h[0] && bwd[k][0] == h[1]; k++) {
        int i0 = F[h[2]][h[3]], i3 = F[bwd[k][2]][bwd[k][3]];
        hinge.push_back({i0, h[0], h[1], i3});
        rest_angle.push_back(hinge_angle(q[i0], q[h[0]], q[h[1]], q[i3]));
        rest_len.push_back(len(q[h[0]] - q[h[1]]));
      }
    }
    hblk.resize(hinge.size());
    if (hinge.size() != E.size()) return -2;   // energy.py:626-628 asserts one hinge per edge
    return 0;
  }

  // geometry_kernels.py:91-104, brute force over the target's triangles
  void update_target() {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < nV; i++) {
      double best = INFINITY;
      V bp = q[i];
      for (auto& t : gtF) {
        V cp = closest_on_tri(q[i], gtV[t[0]], gtV[t[1]], gtV[t[2]]);
        double d = len2(cp - q[i]);
        if (d < best) { best = d; bp = cp; }
      }
      target[i] = bp;
    }
  }

  // contact_detection.py:120-216: pairs within 2 radius + d_hat
  void detect_contact() {
    for (;;) {
      const double thr = 2.0 * radius + c.d_hat;
      con.clear();
      const int nf = (int)F.size(), ne = (int)E.size();
      // sweep along x over boxes padded by thr
      std::vector<std::pair<double, int>> fx(nf), ex(ne);
      std::vector<double> fxh(nf), exh(ne);
      for (int i = 0; i < nf; i++) {
        double lo = std::min(q[F[i][0]].x, std::min(q[F[i][1]].x, q[F[i][2]].x));
        fxh[i] = std::max(q[F[i][0]].x, std::max(q[F[i][1]].x, q[F[i][2]].x));
        fx[i] = {lo, i};
      }
      for (int i = 0; i < ne; i++) {
        ex[i] = {std::min(q[E[i][0]].x, q[E[i][1]].x), i};
        exh[i] = std::max(q[E[i][0]].x, q[E[i][1]].x);
      }
      std::vector<std::pair<double, int>> fxs = fx, exs = ex;
      std::sort(fxs.begin(), fxs.end());
      std::sort(exs.begin(), exs.end());
      std::vector<std::vector<Contact>> per(nV + ne);
      const double pad = thr * (1.0 + 1e-9) + 1e-12;
#pragma omp parallel for schedule(dynamic, 32)
      for (int i0 = 0; i0 < nV; i0++) {
        V x0 = q[i0];
        auto endp = std::upper_bound(fxs.begin(), fxs.end(), std::make_pair(x0.x + pad, INT32_MAX));
        std::vector<std::pair<int, Contact>> hit;   // (triangle, contact)
        for (auto it = fxs.begin(); it != endp; ++it) {
          int t = it->second;
          if (fxh[t] < x0.x - pad) continue;
          int i1 = F[t][0], i2 = F[t][1], i3 = F[t][2];
          if (i0 == i1 || i0 == i2 || i0 == i3) continue;
          int ty = pt_classify(x0, q[i1], q[i2], q[i3]);
          double d = pt_distance(x0, q[i1], q[i2], q[i3], ty);
          if (d <= thr) hit.push_back({t, Contact{BT_PT, {i0, i1, i2, i3}, ty, d, {}}});
        }
        std::sort(hit.begin(), hit.end(),
                  [](const std::pair<int, Contact>& a, const std::pair<int, Contact>& b) { return a.first < b.first; });
        std::vector<Contact> out;
        for (auto& h : hit) out.push_back(h.second);
        per[i0] = out;
      }
#pragma omp parallel for schedule(dynamic, 32)
      for (int e0 = 0; e0 < ne; e0++) {
        int i0 = E[e0][0], i1 = E[e0][1];
        V x0 = q[i0], x1 = q[i1];
        double hi = std::max(x0.x, x1.x) + pad, lo = std::min(x0.x, x1.x) - pad;
        auto endp = std::upper_bound(exs.begin(), exs.end(), std::make_pair(hi, INT32_MAX));
        std::vector<Contact> out;
        for (auto it = exs.begin(); it != endp; ++it) {
          int e1 = it->second;
          if (e1 <= e0 || exh[e1] < lo) continue;
          int i2 = E[e1][0], i3 = E[e1][1];
          if (i0 == i2 || i0 == i3 || i1 == i2 || i1 == i3) continue;
          int ty = ee_classify(x0, x1, q[i2], q[i3], c.ee_thres);
          double d = ee_distance(x0, x1, q[i2], q[i3], ty);
          if (d <= thr) out.push_back(Contact{BT_EE, {i0, i1, i2, i3}, ty, d, {}});
        }
        std::sort(out.begin(), out.end(), [&](const Contact& a, const Contact& b) {
          return std::array<int, 2>{a.i[2], a.i[3]} < std::array<int, 2>{b.i[2], b.i[3]};
        });
        per[nV + e0] = out;
      }
      for (auto& v : per)
        for (auto& x : v) con.push_back(x);
      if ((long long)con.size() > c.max_blocks) {   // energy.py:937-944: retry with half the radius
        radius /= 2.0;
        continue;
      }
      break;
    }
    q_det = q;
  }

  // system.py:343-358: the four energies at x (the contact distances cached)
  double compute_energy(const std::vector<V>& x, double* parts = nullptr) {
    double e[4] = {0, 0, 0, 0};
    for (int i = 0; i < nV; i++) e[0] += 0.5 * len2(x[i] - target[i]) * (vor[i] * c.m2gt_k);
    for (size_t t = 0; t < F.size(); t++)
      e[1] += elastic_energy(x[F[t][0]], x[F[t][1]], x[F[t][2]], W[t], invDm[t], mu, la);
    for (size_t h = 0; h < hinge.size(); h++) {
      double th = hinge_angle(x[hinge[h][0]], x[hinge[h][1]], x[hinge[h][2]], x[hinge[h][3]]);
      double dt = th - rest_angle[h];
      e[2] += 0.5 * c.hinge_k * dt * dt * rest_len[h];
    }
    for (auto& k : con) {
      V a = x[k.i[0]], b = x[k.i[1]], cc = x[k.i[2]], d = x[k.i[3]];
      if (k.type1 == BT_PT) {
        k.type2 = pt_classify(a, b, cc, d);
        k.d = pt_distance(a, b, cc, d, k.type2);
      } else {
        k.type2 = ee_classify(a, b, cc, d, c.ee_thres);
        k.d = ee_distance(a, b, cc, d, k.type2);
      }
      if (k.d < c.d_hat) e[3] += bar(k.d, c.d_hat) * c.kappa;
    }
    if (parts)
      for (int i = 0; i < 4; i++) parts[i] = e[i];
    return e[0] + e[1] + e[2] + e[3];
  }

  // system.py:360-365: grad = -dE/dx, hess_diag; hinge blocks projected after
  void compute_diff() {
    const double coeff = -1.0;
    for (int i = 0; i < nV; i++) {
      double w = vor[i] * c.m2gt_k;
      grad[i] = (q[i] - target[i]) * w * coeff;
      diag[i] = mk(w, w, w);
    }
    for (size_t t = 0; t < F.size(); t++) {
      V f[3];
      elastic_force(q[F[t][0]], q[F[t][1]], q[F[t][2]], W[t], invDm[t], mu, la, coeff, f);
      for (int k = 0; k < 3; k++) grad[F[t][k]] = grad[F[t][k]] + f[k];
    }
    for (size_t h = 0; h < hinge.size(); h++) {
      V g[4];
      M3 blk[4][4];
      const auto& hi = hinge[h];
      hinge_diff(q[hi[0]], q[hi[1]], q[hi[2]], q[hi[3]], rest_angle[h], rest_len[h], c.hinge_k, coeff, g, blk);
      for (int k = 0; k < 4; k++) {
        grad[hi[k]] = grad[hi[k]] + g[k];
        diag[hi[k]] = diag[hi[k]] + mk(blk[k][k].m[0], blk[k][k].m[4], blk[k][k].m[8]);
      }
      spd_project(blk, c.spd_iters);
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) hblk[h][a][b] = blk[a][b];
    }
    for (auto& k : con) {
      for (int j = 0; j < 4; j++) k.dd[j] = mk(0, 0, 0);
      if (k.d >= c.d_hat) continue;
      V x[4] = {q[k.i[0]], q[k.i[1]], q[k.i[2]], q[k.i[3]]};
      if (k.type1 == BT_PT) pt_grad(x[0], x[1], x[2], x[3], k.type2, k.dd);
      else ee_grad(x[0], x[1], x[2], x[3], k.type2, k.dd);
      double g1 = bar1(k.d, c.d_hat) * c.kappa * coeff, g2 = bar2(k.d, c.d_hat) * c.kappa;
      for (int j = 0; j < 4; j++) {
        grad[k.i[j]] = grad[k.i[j]] + k.dd[j] * g1;
        diag[k.i[j]] = diag[k.i[j]] + g2 * cmul(k.dd[j], k.dd[j]);
      }
    }
  }

  // system.py:367-371
  void hess_dx(const std::vector<V>& dx, std::vector<V>& out) {
    out.assign(nV, mk(0, 0, 0));
    for (int i = 0; i < nV; i++) out[i] = dx[i] * (vor[i] * c.m2gt_k);
    for (size_t t = 0; t < F.size(); t++) {
      V f[3];
      const auto& tt = F[t];
      elastic_hdx(q[tt[0]], q[tt[1]], q[tt[2]], dx[tt[0]], dx[tt[1]], dx[tt[2]], W[t], invDm[t], mu, la, f);
      for (int k = 0; k < 3; k++) out[tt[k]] = out[tt[k]] + f[k];
    }
    for (size_t h = 0; h < hinge.size(); h++)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) out[hinge[h][a]] = out[hinge[h][a]] + mv(hblk[h][a][b], dx[hinge[h][b]]);
    for (auto& k : con) {
      if (k.d >= c.d_hat) continue;
      double g2 = bar2(k.d, c.d_hat) * c.kappa;
      double s = 0.0;
      for (int j = 0; j < 4; j++) s += dot(k.dd[j], dx[k.i[j]]);
      for (int j = 0; j < 4; j++) out[k.i[j]] = out[k.i[j]] + k.dd[j] * (s * g2);
    }
  }

  // cg_solver.py:100-165 with the kernels of solver_kernels.py:4-62
  void cg() {
    std::vector<V> r = grad, z(nV), v(nV), Av;
    for (int i = 0; i < nV; i++) {
      p[i] = mk(0, 0, 0);
      z[i] = mk(r[i].x / (diag[i].x + 1e-6), r[i].y / (diag[i].y + 1e-6), r[i].z / (diag[i].z + 1e-6));
    }
    v = z;
    double zr = 0.0;
    for (int i = 0; i < nV; i++) zr += dot(z[i], r[i]);
    for (int it = 0; it < c.n_cg; it++) {
      hess_dx(v, Av);
      double vAv = 0.0;
      for (int i = 0; i < nV; i++) vAv += dot(v[i], Av[i]);
      double zr_last = zr, zr_new = 0.0;
      for (int i = 0; i < nV; i++) {
        if (vAv > 1e-16) {
          double alpha = zr_last / vAv;
          p[i] = p[i] + alpha * v[i];
          r[i] = r[i] - alpha * Av[i];
        }
        z[i] = mk(r[i].x / diag[i].x, r[i].y / diag[i].y, r[i].z / diag[i].z);
        zr_new += dot(z[i], r[i]);
      }
      zr = zr_new;
      double s = zr_new / zr_last;
      for (int i = 0; i < nV; i++) v[i] = z[i] + s * v[i];
    }
  }

  // solver_kernels.py:76-94
  void clamp_p() {
    for (int i = 0; i < nV; i++) {
      V dq = q_newton[i] + p[i] - q_det[i];
      if (len(dq) > c.radius) p[i] = nrm(dq) * c.radius + q_det[i] - q_newton[i];
    }
  }

  void ccd() {
    ccd_step = 1.0;
    for (auto& k : con) {
      V x[4], v[4];
      for (int j = 0; j < 4; j++) { x[j] = q[k.i[j]]; v[j] = p[k.i[j]]; }
      ccd_step = std::min(ccd_step, accd_pair(x, v, k.type1, c.slack, c.thick, c.ccd_iters, c.ee_thres));
    }
  }

  // system.py:395-405 with line_search_kernel (solver_kernels.py:64-74)
  void line_search() {
    energy_prev = energy;
    for (int n = 0; n < c.n_ls; n++) {
      if (!(n > 0 && energy < energy_prev)) {
        double a = ccd_step * std::pow(0.5, (double)n);
        for (int i = 0; i < nV; i++) q[i] = q_newton[i] + a * p[i];
        ls_taken = n;
      }
      energy = compute_energy(q);
    }
  }

  void newton() {
    q_newton = q;
    detect_contact();
    energy = compute_energy(q);
    compute_diff();
    cg();
    clamp_p();
    ccd();
    line_search();
  }

  // system.py:407-436
  void step() {
    if (do_nothing) return;
    update_target();
    for (int it = 0; it < c.n_newton; it++) newton();
  }
};

static Cfg cfg_from(const double* d, const long long* i) {
  Cfg c;
  if (d) {
    c.d_hat = d[0]; c.kappa = d[1]; c.radius = d[2]; c.ee_thres = d[3]; c.slack = d[4]; c.thick = d[5];
    c.young = d[6]; c.poisson = d[7]; c.hinge_k = d[8]; c.m2gt_k = d[9]; c.min_angle = d[10];
  }
  if (i) {
    c.n_newton = (int)i[0]; c.n_cg = (int)i[1]; c.n_ls = (int)i[2]; c.spd_iters = (int)i[3]; c.ccd_iters = (int)i[4];
    c.max_blocks = i[5];
  }
  return c;
}

static inline V ld(const double* p, int k) { return mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]); }
static inline void st(double* p, int k, V v) { p[3 * k] = v.x; p[3 * k + 1] = v.y; p[3 * k + 2] = v.z; }

}  // namespace s3o

using namespace s3o;

extern "C" {

// point-triangle: type, distance and its gradient (12 doubles)
void s3o_pt(const double* x, int* type, double* d, double* g) {
  V a = ld(x, 0), b = ld(x, 1), c = ld(x, 2), e = ld(x, 3);
  int t = pt_classify(a, b, c, e);
  V gg[4];
  pt_grad(a, b, c, e, t, gg);
  *type = t;
  *d = pt_distance(a, b, c, e, t);
  for (int k = 0; k < 4; k++) st(g, k, gg[k]);
}
void s3o_ee(const double* x, double thres, int* type, double* d, double* g) {
  V a = ld(x, 0), b = ld(x, 1), c = ld(x, 2), e = ld(x, 3);
  int t = ee_classify(a, b, c, e, thres);
  V gg[4];
  ee_grad(a, b, c, e, t, gg);
  *type = t;
  *d = ee_distance(a, b, c, e, t);
  for (int k = 0; k < 4; k++) st(g, k, gg[k]);
}
void s3o_closest(const double* p, const double* tri, double* out) {
  st(out, 0, closest_on_tri(ld(p, 0), ld(tri, 0), ld(tri, 1), ld(tri, 2)));
}
void s3o_barrier(double d, double dh, double* out) {
  out[0] = bar(d, dh);
  out[1] = bar1(d, dh);
  out[2] = bar2(d, dh);
}
// rest triangle, current triangle, dx: energy, force (coeff 1: +dpsi/dx), H dx
void s3o_elastic(const double* rest, const double* x, const double* dx, double mu, double la, double* E, double* g,
                 double* hdx) {
  double W;
  M2 inv;
  elastic_rest(ld(rest, 0), ld(rest, 1), ld(rest, 2), W, inv);
  V a = ld(x, 0), b = ld(x, 1), c = ld(x, 2);
  *E = elastic_energy(a, b, c, W, inv, mu, la);
  V f[3];
  elastic_force(a, b, c, W, inv, mu, la, 1.0, f);
  for (int k = 0; k < 3; k++) st(g, k, f[k]);
  elastic_hdx(a, b, c, ld(dx, 0), ld(dx, 1), ld(dx, 2), W, inv, mu, la, f);
  for (int k = 0; k < 3; k++) st(hdx, k, f[k]);
}
// hinge at x with the rest angle and length of rest: angle, energy, +dE/dx,
// 12x12 Hessian (row-major, unprojected) and its projection
void s3o_hinge(const double* rest, const double* x, double k, double* theta, double* E, double* g, double* H,
               double* Hp, int spd_iters) {
  V r0 = ld(rest, 0), r1 = ld(rest, 1), r2 = ld(rest, 2), r3 = ld(rest, 3);
  double ra = hinge_angle(r0, r1, r2, r3), rl = len(r1 - r2);
  V a = ld(x, 0), b = ld(x, 1), c = ld(x, 2), e = ld(x, 3);
  double th = hinge_angle(a, b, c, e);
  *theta = th;
  *E = 0.5 * k * (th - ra) * (th - ra) * rl;
  V gg[4];
  M3 blk[4][4];
  hinge_diff(a, b, c, e, ra, rl, k, 1.0, gg, blk);
  for (int i = 0; i < 4; i++) st(g, i, gg[i]);
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) H[12 * i + j] = blk[i / 3][j / 3].m[3 * (i % 3) + (j % 3)];
  spd_project(blk, spd_iters);
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) Hp[12 * i + j] = blk[i / 3][j / 3].m[3 * (i % 3) + (j % 3)];
}
// projection of an arbitrary 12x12 (only its leading 9x9 is read)
void s3o_spd_project(double* H, int it) {
  M3 blk[4][4];
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) blk[i / 3][j / 3].m[3 * (i % 3) + (j % 3)] = H[12 * i + j];
  spd_project(blk, it);
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) H[12 * i + j] = blk[i / 3][j / 3].m[3 * (i % 3) + (j % 3)];
}
double s3o_accd(const double* x, const double* v, int type1, double slack, double thick, int iters, double ee_thres) {
  V xx[4], vv[4];
  for (int k = 0; k < 4; k++) { xx[k] = ld(x, k); vv[k] = ld(v, k); }
  return accd_pair(xx, vv, type1, slack, thick, iters, ee_thres);
}
void s3o_voronoi(const double* X, int nv, const int* F, int nf, double* out) {
  std::vector<V> x(nv);
  for (int i = 0; i < nv; i++) x[i] = ld(X, i);
  std::vector<std::array<int, 3>> f(nf);
  for (int i = 0; i < nf; i++) f[i] = {F[3 * i], F[3 * i + 1], F[3 * i + 2]};
  std::vector<double> o;
  voronoi(x, f, o);
  std::memcpy(out, o.data(), sizeof(double) * nv);
}

void* s3o_create(const double* cd, const long long* ci) {
  System* s = new System();
  s->setup(cfg_from(cd, ci));
  return s;
}
void s3o_destroy(void* s) { delete (System*)s; }
int s3o_register(void* sp, const double* gv, int ngv, const int* gf, int ngf, const double* v, int nv, const int* f,
                 int nf) {
  return ((System*)sp)->register_mesh(gv, ngv, gf, ngf, v, nv, f, nf);
}
// sizes: nV, nF, nE, n_hinges, n_contacts, do_nothing
void s3o_sizes(void* sp, long long* out) {
  System& s = *(System*)sp;
  out[0] = s.nV; out[1] = (long long)s.F.size(); out[2] = (long long)s.E.size(); out[3] = (long long)s.hinge.size();
  out[4] = (long long)s.con.size(); out[5] = s.do_nothing;
}
void s3o_step(void* sp, int n) {
  for (int i = 0; i < n; i++) ((System*)sp)->step();
}
void s3o_update_target(void* sp) { ((System*)sp)->update_target(); }
// one Newton iteration in parts: 0 detect, 1 energy, 2 diff, 3 cg, 4 clamp, 5 ccd, 6 line search, 7 all
void s3o_newton_part(void* sp, int part) {
  System& s = *(System*)sp;
  switch (part) {
    case 0: s.q_newton = s.q; s.detect_contact(); break;
    case 1: s.energy = s.compute_energy(s.q); break;
    case 2: s.compute_diff(); break;
    case 3: s.cg(); break;
    case 4: s.clamp_p(); break;
    case 5: s.ccd(); break;
    case 6: s.line_search(); break;
    case 7: s.newton(); break;
  }
}
// per-vector state: 0 q, 1 target, 2 grad, 3 diag, 4 p, 5 q_rest, 6 q_newton, 7 q_det (V x 3 doubles)
void s3o_get(void* sp, int which, double* out) {
  System& s = *(System*)sp;
  const std::vector<V>* v[] = {&s.q, &s.target, &s.grad, &s.diag, &s.p, &s.q_rest, &s.q_newton, &s.q_det};
  for (int i = 0; i < s.nV; i++) st(out, i, (*v[which])[i]);
}
void s3o_set(void* sp, int which, const double* in) {
  System& s = *(System*)sp;
  std::vector<V>* v[] = {&s.q, &s.target, &s.grad, &s.diag, &s.p, &s.q_rest, &s.q_newton, &s.q_det};
  for (int i = 0; i < s.nV; i++) (*v[which])[i] = ld(in, i);
}
// scalars: energy, energy_prev, ccd_step, radius, line-search halvings taken
void s3o_scalars(void* sp, double* out) {
  System& s = *(System*)sp;
  out[0] = s.energy; out[1] = s.energy_prev; out[2] = s.ccd_step; out[3] = s.radius; out[4] = (double)s.ls_taken;
}
void s3o_energy_parts(void* sp, const double* x, double* parts) {
  System& s = *(System*)sp;
  std::vector<V> xx(s.nV);
  for (int i = 0; i < s.nV; i++) xx[i] = ld(x, i);
  s.compute_energy(xx, parts);
}
// contacts: rows of (type1, i0, i1, i2, i3, type2) and the distances
void s3o_contacts(void* sp, int* rows, double* d) {
  System& s = *(System*)sp;
  for (size_t k = 0; k < s.con.size(); k++) {
    const Contact& c = s.con[k];
    int* r = rows + 6 * k;
    r[0] = c.type1; r[1] = c.i[0]; r[2] = c.i[1]; r[3] = c.i[2]; r[4] = c.i[3]; r[5] = c.type2;
    if (d) d[k] = c.d;
  }
}
void s3o_hess_dx(void* sp, const double* dx, double* out) {
  System& s = *(System*)sp;
  std::vector<V> d(s.nV), o;
  for (int i = 0; i < s.nV; i++) d[i] = ld(dx, i);
  s.hess_dx(d, o);
  for (int i = 0; i < s.nV; i++) st(out, i, o[i]);
}
void s3o_edges(void* sp, int* out) {
  System& s = *(System*)sp;
  for (size_t i = 0; i < s.E.size(); i++) { out[2 * i] = s.E[i][0]; out[2 * i + 1] = s.E[i][1]; }
}
void s3o_hinges(void* sp, int* idx, double* rest) {
  System& s = *(System*)sp;
  for (size_t i = 0; i < s.hinge.size(); i++) {
    for (int k = 0; k < 4; k++) idx[4 * i + k] = s.hinge[i][k];
    rest[2 * i] = s.rest_angle[i];
    rest[2 * i + 1] = s.rest_len[i];
  }
}
void s3o_voronoi_of(void* sp, double* out) {
  System& s = *(System*)sp;
  for (int i = 0; i < s.nV; i++) out[i] = s.vor[i];
}

}  // extern "C"
