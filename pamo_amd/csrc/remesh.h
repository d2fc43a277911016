// Stage 1 (remeshing) arithmetic shared by the device kernels and the host
// oracle: point-triangle distance, and the cube polygonisation of the
// extraction.  The reference calls two external packages here
// (simp_cuda/pamo/__init__.py:75-93): torchcumesh2sdf.get_sdf (a narrow-band
// signed distance volume of a triangle soup) and pdmc.DMC (its zero level
// set as a mesh); neither ships with the reference.  This restates what the
// call sites need from them: a signed distance on an R^3 lattice, exact
// within the band, whose sign comes from a flood fill of the outside, and a
// closed, consistently oriented 2-manifold through the zero crossings.
//
// Every operation is a single IEEE operation in the order written, so the
// host oracle (g++ -ffp-contract=off) and the device (hipcc
// -ffp-contract=off) compute the same bits.
#ifndef PAMO_REMESH_H
#define PAMO_REMESH_H

#include <stdint.h>

#ifdef __HIPCC__
#define RM_FN __host__ __device__ inline
#else
#define RM_FN inline
#endif

namespace pamo {
namespace rm {

struct F3 { float x, y, z; };
RM_FN F3 f3sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RM_FN F3 f3add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RM_FN F3 f3mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RM_FN float f3dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RM_FN F3 f3cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Closest point of segment ab to p (clamped projection).
RM_FN F3 closest_on_segment(F3 p, F3 a, F3 b) {
  const F3 ab = f3sub(b, a);
  const float l2 = f3dot(ab, ab);
  if (!(l2 > 0.0f)) return a;
  float t = f3dot(f3sub(p, a), ab) / l2;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  return f3add(a, f3mul(ab, t));
}

// Closest point of triangle abc to p, by the Voronoi regions of its
// vertices, edges and face (C. Ericson, Real-Time Collision Detection, 2005,
// section 5.1.5).  A degenerate triangle (no face region) takes the nearest
// of its three edges' closest points.
RM_FN F3 closest_on_triangle(F3 p, F3 a, F3 b, F3 c) {
  const F3 ab = f3sub(b, a), ac = f3sub(c, a), ap = f3sub(p, a);
  const float d1 = f3dot(ab, ap), d2 = f3dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const F3 bp = f3sub(p, b);
  const float d3 = f3dot(ab, bp), d4 = f3dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return f3add(a, f3mul(ab, d1 / (d1 - d3)));
  const F3 cp = f3sub(p, c);
  const float d5 = f3dot(ab, cp), d6 = f3dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return f3add(a, f3mul(ac, d2 / (d2 - d6)));
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    return f3add(b, f3mul(f3sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  const float den = va + vb + vc;
  if (!(den > 0.0f)) {   // degenerate: the nearest edge point
    const F3 q0 = closest_on_segment(p, a, b), q1 = closest_on_segment(p, b, c), q2 = closest_on_segment(p, c, a);
    const float e0 = f3dot(f3sub(p, q0), f3sub(p, q0)), e1 = f3dot(f3sub(p, q1), f3sub(p, q1)),
                e2 = f3dot(f3sub(p, q2), f3sub(p, q2));
    return e0 <= e1 ? (e0 <= e2 ? q0 : q2) : (e1 <= e2 ? q1 : q2);
  }
  const float v = vb / den, w = vc / den;
  return f3add(a, f3add(f3mul(ab, v), f3mul(ac, w)));
}

// the lattice: R^3 samples, sample (i, j, k) at ((i + 0.5) / R, ...), index
// (k R + j) R + i
RM_FN float sample_coord(int i, int R) { return ((float)i + 0.5f) / (float)R; }

// The distance key of a sample: the distance's bits (non-negative floats
// order as their bits) over the triangle index, so the minimum over
// triangles is one 64-bit min and ties go to the lower triangle.
RM_FN unsigned long long dist_key(float d, int tri) {
  union { float f; uint32_t u; } x;
  x.f = d;
  return ((unsigned long long)x.u << 32) | (uint32_t)tri;
}
RM_FN float key_dist(unsigned long long k) {
  union { float f; uint32_t u; } x;
  x.u = (uint32_t)(k >> 32);
  return x.f;
}
constexpr unsigned long long NO_KEY = ~0ull;

// sample states of the sign pass
constexpr uint8_t ST_PASS = 1;   // the flood may enter: no surface within tau_p
constexpr uint8_t ST_OUT = 2;    // reached from the lattice boundary (outside)

// The side of a sample relative to its nearest triangle: the sign of
// (p - closest point) . (b - a) x (c - a); 0 when on the plane.
RM_FN int tri_side(F3 p, F3 a, F3 b, F3 c) {
  const F3 q = closest_on_triangle(p, a, b, c);
  const float s = f3dot(f3sub(p, q), f3cross(f3sub(b, a), f3sub(c, a)));
  return s > 0.0f ? 1 : (s < 0.0f ? -1 : 0);
}

// --------------------------------------------------------- polygonisation
// Cube: corner c = dx + 2 dy + 4 dz; edges 0-3 along x (e = y + 2z), 4-7
// along y (e = 4 + x + 2z), 8-11 along z (e = 8 + x + 2y); faces 2a + s lie
// at coordinate a = s, their corners in cyclic order, the edges between
// consecutive corners, the outward normal.
struct McTab {
  int ec[12][2];
  int fc[6][4];
  int fe[6][4];
  int fn[6][3];
};
constexpr McTab mc_tables() {
  McTab t{};
  for (int y = 0; y < 2; y++)
    for (int z = 0; z < 2; z++) {
      t.ec[y + 2 * z][0] = 2 * y + 4 * z;
      t.ec[y + 2 * z][1] = 1 + 2 * y + 4 * z;
    }
  for (int x = 0; x < 2; x++)
    for (int z = 0; z < 2; z++) {
      t.ec[4 + x + 2 * z][0] = x + 4 * z;
      t.ec[4 + x + 2 * z][1] = x + 2 + 4 * z;
    }
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++) {
      t.ec[8 + x + 2 * y][0] = x + 2 * y;
      t.ec[8 + x + 2 * y][1] = x + 2 * y + 4;
    }
  const int cyc[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  for (int a = 0; a < 3; a++)
    for (int s = 0; s < 2; s++) {
      const int f = 2 * a + s, b = (a + 1) % 3, c = (a + 2) % 3;
      for (int q = 0; q < 4; q++) {
        int p[3] = {0, 0, 0};
        p[a] = s;
        p[b] = cyc[q][0];
        p[c] = cyc[q][1];
        t.fc[f][q] = p[0] + 2 * p[1] + 4 * p[2];
      }
      for (int q = 0; q < 4; q++) {
        const int u = t.fc[f][q], w = t.fc[f][(q + 1) % 4];
        for (int e = 0; e < 12; e++)
          if ((t.ec[e][0] == u && t.ec[e][1] == w) || (t.ec[e][0] == w && t.ec[e][1] == u)) t.fe[f][q] = e;
      }
      t.fn[f][a] = s ? 1 : -1;
    }
  return t;
}

// The surface pieces of one cube as directed segments on its faces: on a
// face, the crossing points (on the edges whose ends differ in sign, > 0
// against <= 0) are joined in pairs; a face with four crossings joins
// across the diagonal whose value product is larger (ties: the positive
// diagonal), which both cubes sharing the face decide alike.  Each segment
// is directed so that, seen from outside the cube, the positive side lies
// to its left; chained, the segments close into loops whose normal points
// to the positive side.  next[e] is the edge that follows e (-1: none).
// Returns false on an inconsistency (an edge with two successors).
RM_FN long long det3i(const long long a[3], const long long b[3], const long long c[3]) {
  return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}
RM_FN void corner_pos(int c, long long p[3]) {
  p[0] = c & 1;
  p[1] = (c >> 1) & 1;
  p[2] = (c >> 2) & 1;
}
RM_FN bool seg_add(const McTab& T, int f, int eA, int eB, const long long nu[3], int next[12]) {
  long long pa[3], qa[3], pb[3], qb[3];
  corner_pos(T.ec[eA][0], pa);
  corner_pos(T.ec[eA][1], qa);
  corner_pos(T.ec[eB][0], pb);
  corner_pos(T.ec[eB][1], qb);
  long long d[3], n[3];
  for (int k = 0; k < 3; k++) {
    d[k] = (pb[k] + qb[k]) - (pa[k] + qa[k]);
    n[k] = T.fn[f][k];
  }
  const bool fwd = det3i(d, nu, n) > 0;
  const int a = fwd ? eA : eB, b = fwd ? eB : eA;
  if (next[a] != -1) return false;
  next[a] = b;
  return true;
}
// A cube's code: bit c (0..7) corner c positive; bit 8 + f face f joins its
// 0-2 diagonal (set only on faces with four crossings).  The segments, and
// so the triangles, depend on the code alone.
RM_FN int cube_code(const McTab& T, const float val[8]) {
  int code = 0;
  for (int c = 0; c < 8; c++)
    if (val[c] > 0.0f) code |= 1 << c;
  for (int f = 0; f < 6; f++) {
    bool pos[4];
    for (int q = 0; q < 4; q++) pos[q] = val[T.fc[f][q]] > 0.0f;
    if (pos[0] == pos[1] || pos[1] == pos[2] || pos[2] == pos[3]) continue;   // not four crossings
    const float p02 = val[T.fc[f][0]] * val[T.fc[f][2]], p13 = val[T.fc[f][1]] * val[T.fc[f][3]];
    if (p02 > p13 || (p02 == p13 && pos[0])) code |= 1 << (8 + f);
  }
  return code;
}
RM_FN bool cube_segments_code(const McTab& T, int code, int next[12]) {
  for (int e = 0; e < 12; e++) next[e] = -1;
  for (int f = 0; f < 6; f++) {
    bool pos[4];
    for (int q = 0; q < 4; q++) pos[q] = (code >> T.fc[f][q]) & 1;
    int ncross = 0;
    for (int q = 0; q < 4; q++) ncross += pos[q] != pos[(q + 1) % 4];
    if (ncross == 0) continue;
    long long P[4][3];
    for (int q = 0; q < 4; q++) corner_pos(T.fc[f][q], P[q]);
    if (ncross == 2) {
      int es[2] = {0, 0}, m = 0;
      for (int q = 0; q < 4; q++)
        if (pos[q] != pos[(q + 1) % 4]) es[m++] = T.fe[f][q];
      // from the negative corners' centroid towards the positive ones'
      long long sp[3] = {0, 0, 0}, sn[3] = {0, 0, 0};
      int np = 0, nn = 0;
      for (int q = 0; q < 4; q++) {
        for (int k = 0; k < 3; k++) {
          if (pos[q]) sp[k] += P[q][k];
          else sn[k] += P[q][k];
        }
        if (pos[q]) np++;
        else nn++;
      }
      long long nu[3];
      for (int k = 0; k < 3; k++) nu[k] = sp[k] * (12 / np) - sn[k] * (12 / nn);
      if (!seg_add(T, f, es[0], es[1], nu, next)) return false;
    } else {
      const bool conn02 = (code >> (8 + f)) & 1;
      const int cuts[2] = {conn02 ? 1 : 0, conn02 ? 3 : 2};
      for (int ci = 0; ci < 2; ci++) {
        const int q = cuts[ci];
        const int eA = T.fe[f][(q + 3) % 4], eB = T.fe[f][q];
        long long sum[3] = {0, 0, 0};
        for (int r = 0; r < 4; r++)
          for (int k = 0; k < 3; k++) sum[k] += P[r][k];
        long long nu[3];
        for (int k = 0; k < 3; k++) nu[k] = pos[q] ? (4 * P[q][k] - sum[k]) : (sum[k] - 4 * P[q][k]);
        if (!seg_add(T, f, eA, eB, nu, next)) return false;
      }
    }
  }
  return true;
}
RM_FN bool cube_segments(const McTab& T, const float val[8], int next[12]) {
  return cube_segments_code(T, cube_code(T, val), next);
}
constexpr int CUBE_CODES = 1 << 14;
// The cube's triangles.  Every loop of next[] (taken from its lowest edge)
// is fanned from the first of its edges whose fan has no triangle lying flat
// in a cube face (all three corners' edges on one face: on a saddle face the
// neighbouring cube would emit the same triangle reversed); a loop without
// such an apex is fanned from a centre vertex instead, the mean of its
// crossing points.  tris[] receives local corners (a cube edge 0..11, or
// 12 + x for the cube's x-th centre vertex), centres[x] the loop of centre x
// (edges in loop order, terminated by -1 unless 12 long).  Returns the
// triangle count (centres in *ncent), or -1 for a loop that does not close.
struct CubeOut {
  int tris[12][3];
  int cent[4][13];
};
RM_FN int cube_tris(const McTab& T, const int next[12], CubeOut* out, int* ncent) {
  int fmask[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int f = 0; f < 6; f++)
    for (int q = 0; q < 4; q++) fmask[T.fe[f][q]] |= 1 << f;
  bool seen[12] = {false, false, false, false, false, false, false, false, false, false, false, false};
  int nt = 0, nc = 0;
  for (int e0 = 0; e0 < 12; e0++) {
    if (next[e0] < 0 || seen[e0]) continue;
    int loop[12], m = 0, e = e0;
    while (e >= 0 && !seen[e] && m < 12) {
      seen[e] = true;
      loop[m++] = e;
      e = next[e];
    }
    if (e != e0 || m < 3) return -1;
    int apex = -1;
    for (int r = 0; r < m && apex < 0; r++) {
      bool flat = false;
      for (int q = 1; q + 1 < m && !flat; q++)
        flat = (fmask[loop[r]] & fmask[loop[(r + q) % m]] & fmask[loop[(r + q + 1) % m]]) != 0;
      if (!flat) apex = r;
    }
    if (apex >= 0) {
      for (int q = 1; q + 1 < m; q++) {
        if (out) {
          out->tris[nt][0] = loop[apex];
          out->tris[nt][1] = loop[(apex + q) % m];
          out->tris[nt][2] = loop[(apex + q + 1) % m];
        }
        nt++;
      }
    } else {
      if (out) {
        for (int q = 0; q < m; q++) out->cent[nc][q] = loop[q];
        if (m < 13) out->cent[nc][m] = -1;
      }
      for (int q = 0; q < m; q++) {
        if (out) {
          out->tris[nt][0] = 12 + nc;
          out->tris[nt][1] = loop[q];
          out->tris[nt][2] = loop[(q + 1) % m];
        }
        nt++;
      }
      nc++;
    }
  }
  *ncent = nc;
  return nt;
}
// the lattice edge of a cube edge: 3 (first sample) + axis
RM_FN long long cube_edge_id(const McTab& T, int e, int i, int j, int k, int R) {
  const int c = T.ec[e][0];
  const long long s = ((long long)(k + ((c >> 2) & 1)) * R + (j + ((c >> 1) & 1))) * R + (i + (c & 1));
  return 3 * s + (e < 4 ? 0 : (e < 8 ? 1 : 2));
}
// the crossing of lattice edge (sample s, axis) at t in [0.001, 0.999] of
// the way from s to its neighbour (values a at s, b at the neighbour)
RM_FN float crossing_t(float a, float b) {
  float t = a / (a - b);
  return t < 0.001f ? 0.001f : (t > 0.999f ? 0.999f : t);
}

}  // namespace rm
}  // namespace pamo

#endif
