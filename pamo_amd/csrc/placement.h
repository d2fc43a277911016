// Optimal-vertex placement for an edge collapse: an opt-in mode with no
// counterpart in the reference, which scores and places every collapse at
// the edge midpoint (simp_cuda/src/cusimp_free.cu:491,687-688).
//
// Q = Q_u + Q_v is the summed quadric, given as its 10 unique entries
//   q0 m00, q1 m01, q2 m02, q3 m03, q4 m11, q5 m12, q6 m13, q7 m22, q8 m23,
//   q9 m33.
// The placement minimises v^T Q v over v = (x, 1): A x = -b with A the upper
// 3x3 block and b = (m03, m13, m23).
//   1. FP32: Cramer's rule when det(A) > 1e-3 tr(A)^3 (A is a sum of n n^T
//      over unit face normals, so det(A) <= (tr(A)/3)^3: the test bounds the
//      condition number).
//   2. FP64: otherwise the same cofactors and solve in double from the same
//      float entries, when det(A) > 1e-10 tr(A)^3; the solution is rounded
//      to float once.
//   3. A solution farther from the edge midpoint than the edge is long (or
//      NaN) is rejected.
//   4. Otherwise, or after a rejection, the midpoint / endpoint rule: the
//      one of the midpoint, u and v with the lowest v^T Q v (ties in that
//      order).
// Every operation is a single IEEE operation in the order written, so the
// host oracle (g++ -ffp-contract=off) and the device (hipcc
// -ffp-contract=off) compute the same bits.
#ifndef PAMO_PLACEMENT_H
#define PAMO_PLACEMENT_H

#ifdef __HIPCC__
#define PL_FN __host__ __device__ inline
#else
#define PL_FN inline
#endif

namespace pamo {
namespace place {

struct P3 { float x, y, z; };

enum How { MIDPOINT_REF = 0, SOLVE_F32 = 1, SOLVE_F64 = 2, RULE_MID = 3, RULE_U = 4, RULE_V = 5 };

// v^T Q v at (p, 1), in the order of the cost kernel (cusimp_free.h:145-152)
PL_FN float qem(const float* q, P3 p) {
  const float t0 = q[0] * p.x + q[1] * p.y + q[2] * p.z + q[3] * 1.0f;
  const float t1 = q[1] * p.x + q[4] * p.y + q[5] * p.z + q[6] * 1.0f;
  const float t2 = q[2] * p.x + q[5] * p.y + q[7] * p.z + q[8] * 1.0f;
  const float t3 = q[3] * p.x + q[6] * p.y + q[8] * p.z + q[9] * 1.0f;
  return t0 * p.x + t1 * p.y + t2 * p.z + t3 * 1.0f;
}

PL_FN P3 midpoint(P3 a, P3 b) { return {(a.x + b.x) / 2.0f, (a.y + b.y) / 2.0f, (a.z + b.z) / 2.0f}; }

// The placement of the collapse of edge (a, b) under quadric q; *how names
// the branch taken.
PL_FN P3 optimal(const float* q, P3 a, P3 b, int* how) {
  const P3 m = midpoint(a, b);
  const float tr = q[0] + q[4] + q[7];
  P3 x = m;
  int h = -1;
  if (tr > 0.0f) {
    const float c00 = q[4] * q[7] - q[5] * q[5];
    const float c01 = q[2] * q[5] - q[1] * q[7];
    const float c02 = q[1] * q[5] - q[2] * q[4];
    const float det = q[0] * c00 + q[1] * c01 + q[2] * c02;
    if (det > 1e-3f * tr * tr * tr) {
      const float c11 = q[0] * q[7] - q[2] * q[2];
      const float c12 = q[1] * q[2] - q[0] * q[5];
      const float c22 = q[0] * q[4] - q[1] * q[1];
      x.x = -(c00 * q[3] + c01 * q[6] + c02 * q[8]) / det;
      x.y = -(c01 * q[3] + c11 * q[6] + c12 * q[8]) / det;
      x.z = -(c02 * q[3] + c12 * q[6] + c22 * q[8]) / det;
      h = SOLVE_F32;
    } else {
      double d[10];
      for (int i = 0; i < 10; i++) d[i] = (double)q[i];
      const double e00 = d[4] * d[7] - d[5] * d[5];
      const double e01 = d[2] * d[5] - d[1] * d[7];
      const double e02 = d[1] * d[5] - d[2] * d[4];
      const double dd = d[0] * e00 + d[1] * e01 + d[2] * e02;
      const double trd = (double)tr;
      if (dd > 1e-10 * trd * trd * trd) {
        const double e11 = d[0] * d[7] - d[2] * d[2];
        const double e12 = d[1] * d[2] - d[0] * d[5];
        const double e22 = d[0] * d[4] - d[1] * d[1];
        x.x = (float)(-(e00 * d[3] + e01 * d[6] + e02 * d[8]) / dd);
        x.y = (float)(-(e01 * d[3] + e11 * d[6] + e12 * d[8]) / dd);
        x.z = (float)(-(e02 * d[3] + e12 * d[6] + e22 * d[8]) / dd);
        h = SOLVE_F64;
      }
    }
    if (h >= 0) {
      const float dx = x.x - m.x, dy = x.y - m.y, dz = x.z - m.z;
      const float lx = a.x - b.x, ly = a.y - b.y, lz = a.z - b.z;
      if (!(dx * dx + dy * dy + dz * dz <= lx * lx + ly * ly + lz * lz)) h = -1;
    }
  }
  if (h < 0) {
    const float em = qem(q, m), ea = qem(q, a), eb = qem(q, b);
    x = m;
    h = RULE_MID;
    float best = em;
    if (ea < best) { x = a; best = ea; h = RULE_U; }
    if (eb < best) { x = b; h = RULE_V; }
  }
  if (how) *how = h;
  return x;
}

}  // namespace place
}  // namespace pamo

#endif
