// The flip test of an edge collapse, decided from the unnormalised normal of
// the moved face where the answer is provable, without the square root and
// the three divisions of the normalisation.
//
// The reference marks a collapse as turning a face over when
//   dot(a, b / |b|) < 0                        (cusimp_free.cu:520-526,560-566)
// with a the unit normal of the face before the collapse, b the cross
// product of the moved face, every operation rounded to FP32 with no FMA
// contraction: L = fl(sqrt(fl(b.b))), q_i = fl(b_i / L),
// r = fl(fl(a0 q0 + a1 q1) + a2 q2).
//
// flip_decide(a, b) returns 1 (r < 0), 0 (not r < 0) or -1 (not decided:
// the caller evaluates the expression above).  It decides when
//   (i)   max |a_i| <= 2,
//   (ii)  B = fl(|b0| + |b1| + |b2|) lies in [2^-40, 2^40],
//   (iii) |d| > 2^-20 Tc and |d| > 2^-80 B, with d = fl-dot(a, b) and
//         Tc = fl(|fl(a0 b0)| + |fl(a1 b1)| + |fl(a2 b2)|),
// and then answers d < 0.  NaN or infinite inputs fail (i)-(iii).
//
// Why (u = 2^-24, S = sum a_i b_i and T = sum |a_i b_i| exactly):
//  * By (ii) fl(b.b) neither overflows nor falls below the normal range, so
//    L is finite with L >= |b| (1 - 2u) > 2^-41.
//  * q_i = (b_i / L)(1 + d_i) + e_i and the products likewise, |d_i| <= u,
//    |e_i| <= 2^-150 (underflow), and the two sums add relative errors
//    <= u; with (i): |r - S/L| <= 4.01 u T/L + 9.1 * 2^-150.
//  * |d - S| <= 3.01 u T + 3 * 2^-150 and Tc >= T (1 - 3.01u) - 3.01 * 2^-150,
//    so (iii) gives |S| > 12.98 u T - 3.01 * 2^-150 and
//    |S| > 0.9999 * 2^-80 L - 3.01 u T - 3 * 2^-150.
//  * If 4.01 u T >= 2^-82 L, then T >= 2^-60 L and the first bound exceeds
//    4.01 u T + 9.1 * 2^-150 L; otherwise the second does (L > 2^-41).
//    Either way |S| / L > |r - S/L|: r is non-zero with the sign of S.
//  * |d| > max(16 u Tc, 2^-120) exceeds |d - S|, so d has the sign of S.
// Hence r < 0 exactly when d < 0.  The thresholds are rounded products;
// their rounding (2^-150 at most) is inside the margins above.
#ifndef PAMO_FLIP_FILTER_H
#define PAMO_FLIP_FILTER_H

#ifdef __HIPCC__
#define FF_FN __host__ __device__ inline
#else
#include <math.h>
#define FF_FN inline
#endif

namespace pamo {
namespace flipf {

// 1: flips, 0: does not, -1: undecided.  (ax, ay, az) the old unit normal,
// (bx, by, bz) the cross product of the moved face.
FF_FN int flip_decide(float ax, float ay, float az, float bx, float by, float bz) {
  const float p0 = ax * bx, p1 = ay * by, p2 = az * bz;
  const float d = p0 + p1 + p2;
  const float tc = fabsf(p0) + fabsf(p1) + fabsf(p2);
  const float bsum = fabsf(bx) + fabsf(by) + fabsf(bz);
  const float ad = fabsf(d);
  const bool ok = fabsf(ax) <= 2.0f && fabsf(ay) <= 2.0f && fabsf(az) <= 2.0f &&
                  bsum >= 0x1p-40f && bsum <= 0x1p40f && ad > 0x1p-20f * tc && ad > 0x1p-80f * bsum;
  return ok ? (d < 0.0f ? 1 : 0) : -1;
}

}  // namespace flipf
}  // namespace pamo

#endif
