// Stage 1 (remeshing) kernels: the signed distance volume of a triangle
// soup and the polygonisation of its zero level set (arithmetic in
// remesh.h).  Call sites: simp_cuda/pamo/__init__.py:75-93 (get_sdf, then
// the volume shifted by the caller, then DMC).
#pragma once
#include "remesh.h"

namespace pamo {
namespace rmk {

using rm::F3;
__constant__ rm::McTab kMc = rm::mc_tables();

__device__ __forceinline__ F3 tri_corner(const float* __restrict__ tris, long long t, int c) {
  const float* q = tris + 9 * t + 3 * c;
  return F3{q[0], q[1], q[2]};
}
__device__ __forceinline__ F3 sample_pos(int i, int j, int k, int R) {
  return F3{rm::sample_coord(i, R), rm::sample_coord(j, R), rm::sample_coord(k, R)};
}
// sample range [lo, hi] of one axis whose coordinates lie within band of
// [a, b] (the same float expressions as the oracle's)
__device__ __forceinline__ void axis_range(float a, float b, float band, int R, int& lo, int& hi) {
  lo = (int)ceilf((a - band) * (float)R - 0.5f);
  hi = (int)floorf((b + band) * (float)R - 0.5f);
  lo = lo < 0 ? 0 : lo;
  hi = hi > R - 1 ? R - 1 : hi;
}

// Unsigned distance within the band: each workgroup takes triangles in
// turn, its lanes the samples of the triangle's box grown by the band; a
// sample keeps the minimum of (distance, triangle) over every triangle
// within the band (dist_key), so the result does not depend on the order.
__global__ void __launch_bounds__(BLOCK) k_rm_udf(const float* __restrict__ tris, long long F, int R, float band,
                                                  unsigned long long* __restrict__ key) {
  for (long long t = blockIdx.x; t < F; t += gridDim.x) {   // workgroup-uniform
    const F3 a = tri_corner(tris, t, 0), b = tri_corner(tris, t, 1), c = tri_corner(tris, t, 2);
    int i0, i1, j0, j1, k0, k1;
    axis_range(fminf(a.x, fminf(b.x, c.x)), fmaxf(a.x, fmaxf(b.x, c.x)), band, R, i0, i1);
    axis_range(fminf(a.y, fminf(b.y, c.y)), fmaxf(a.y, fmaxf(b.y, c.y)), band, R, j0, j1);
    axis_range(fminf(a.z, fminf(b.z, c.z)), fmaxf(a.z, fmaxf(b.z, c.z)), band, R, k0, k1);
    if (i1 < i0 || j1 < j0 || k1 < k0) continue;
    // (a box holds at most R^3 <= 2^27 samples: 32-bit index arithmetic)
    const unsigned ni = i1 - i0 + 1, nj = j1 - j0 + 1, nij = ni * nj, n = nij * (unsigned)(k1 - k0 + 1);
    for (unsigned q = threadIdx.x; q < n; q += blockDim.x) {
      const unsigned qk = q / nij, qr = q - qk * nij, qj = qr / ni;
      const int i = i0 + (int)(qr - qj * ni), j = j0 + (int)qj, k = k0 + (int)qk;
      const F3 p = sample_pos(i, j, k, R);
      const F3 d = rm::f3sub(p, rm::closest_on_triangle(p, a, b, c));
      const float dist = sqrtf(rm::f3dot(d, d));
      if (dist <= band) atomicMin(&key[((long long)k * R + j) * R + i], rm::dist_key(dist, (int)t));
    }
  }
}
// passable samples (no surface within tau_p) and the boundary seeds
__global__ void k_rm_state(const unsigned long long* __restrict__ key, int R, float tau_p, uint8_t* __restrict__ st) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)R * R * R;
  if (s >= n) return;
  const int i = (int)(s % R), j = (int)((s / R) % R), k = (int)(s / ((long long)R * R));
  const unsigned long long q = key[s];
  const bool pass = q == rm::NO_KEY || rm::key_dist(q) >= tau_p;
  const bool edge = i == 0 || j == 0 || k == 0 || i == R - 1 || j == R - 1 || k == R - 1;
  st[s] = (uint8_t)((pass ? rm::ST_PASS : 0) | (pass && edge ? rm::ST_OUT : 0));
}
// One directional sweep of the flood: a lane per lattice line along axis,
// walked forwards (dir 0) or backwards; the outside runs on through
// passable samples.  Repeated over the six directions until nothing
// changes, it marks every passable sample connected to the boundary.
__global__ void k_rm_sweep(uint8_t* __restrict__ st, int R, int axis, int dir, int* __restrict__ changed) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (long long)R * R) return;
  const int u = (int)(l % R), w = (int)(l / R);   // the other two coordinates, fastest first
  const long long RR = (long long)R * R;
  long long base, step;
  if (axis == 0) { base = ((long long)w * R + u) * R; step = 1; }
  else if (axis == 1) { base = (long long)w * RR + u; step = R; }
  else { base = (long long)w * R + u; step = RR; }
  if (dir) { base += (long long)(R - 1) * step; step = -step; }
  bool carry = false, ch = false;
  long long s = base;
  for (int x = 0; x < R; x++, s += step) {
    const uint8_t v = st[s];
    if (!(v & rm::ST_PASS)) { carry = false; continue; }
    if (v & rm::ST_OUT) carry = true;
    else if (carry) { st[s] = v | rm::ST_OUT; ch = true; }
  }
  if (ch) *changed = 1;
}
// A triangle's orientation by vote of the passable samples nearest to it:
// an outside sample on its normal's side, or an inside one on the other,
// says the normal points out.
__global__ void k_rm_votes(const unsigned long long* __restrict__ key, const uint8_t* __restrict__ st,
                           const float* __restrict__ tris, int R, int* __restrict__ votes) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (long long)R * R * R) return;
  const unsigned long long q = key[s];
  const uint8_t v = st[s];
  if (q == rm::NO_KEY || !(v & rm::ST_PASS)) return;
  const long long t = (long long)(uint32_t)q;
  const int i = (int)(s % R), j = (int)((s / R) % R), k = (int)(s / ((long long)R * R));
  const int side = rm::tri_side(sample_pos(i, j, k, R), tri_corner(tris, t, 0), tri_corner(tris, t, 1),
                                tri_corner(tris, t, 2));
  if (side == 0) return;
  atomicAdd(&votes[t], ((v & rm::ST_OUT) != 0) == (side > 0) ? 1 : -1);
}
// The signed distance: beyond the band +-band by the flood; passable
// samples by the flood; samples within tau_p of the surface by the side of
// their nearest triangle and its orientation.
__global__ void k_rm_final(const unsigned long long* __restrict__ key, const uint8_t* __restrict__ st,
                           const float* __restrict__ tris, const int* __restrict__ votes, int R, float band,
                           float* __restrict__ vol) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (long long)R * R * R) return;
  const unsigned long long q = key[s];
  const uint8_t v = st[s];
  bool out;
  float u;
  if (q == rm::NO_KEY) {
    out = (v & rm::ST_OUT) != 0;
    u = band;
  } else {
    u = rm::key_dist(q);
    if (v & rm::ST_PASS) {
      out = (v & rm::ST_OUT) != 0;
    } else {
      const long long t = (long long)(uint32_t)q;
      const int i = (int)(s % R), j = (int)((s / R) % R), k = (int)(s / ((long long)R * R));
      const int side = rm::tri_side(sample_pos(i, j, k, R), tri_corner(tris, t, 0), tri_corner(tris, t, 1),
                                    tri_corner(tris, t, 2));
      out = side * (votes[t] >= 0 ? 1 : -1) > 0;
    }
  }
  vol[s] = out ? u : -u;
}

// ------------------------------------------------------------ extraction
// lattice edge e = 3 s + axis: its crossing flag (sign change, > 0 against
// <= 0), scanned to the vertex ids
struct RmEdgeOp {
  const float* vol;
  int R;
  int* ids;
  struct Rec { int c; };
  __device__ Rec load(long long e) const {
    const long long s = e / 3;
    const int axis = (int)(e % 3);
    const int i = (int)(s % R), j = (int)((s / R) % R), k = (int)(s / ((long long)R * R));
    const int x = axis == 0 ? i : (axis == 1 ? j : k);
    Rec r;
    r.c = 0;
    if (x < R - 1) {
      const long long o = axis == 0 ? 1 : (axis == 1 ? R : (long long)R * R);
      r.c = (vol[s] > 0.0f) != (vol[s + o] > 0.0f) ? 1 : 0;
    }
    return r;
  }
  __device__ int count(const Rec& r) const { return r.c; }
  __device__ void gather(long long, Rec&) const {}
  __device__ void emit(long long e, const Rec& r, int pos) const { ids[e] = r.c ? pos : -1; }
};
// crossing positions, in the lattice's [0, 1] coordinates
__global__ void k_rm_vpos(const float* __restrict__ vol, int R, const int* __restrict__ ids,
                          float* __restrict__ verts) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3LL * R * R * R) return;
  const int id = ids[e];
  if (id < 0) return;
  const long long s = e / 3;
  const int axis = (int)(e % 3);
  const int i = (int)(s % R), j = (int)((s / R) % R), k = (int)(s / ((long long)R * R));
  const long long o = axis == 0 ? 1 : (axis == 1 ? R : (long long)R * R);
  const float t = rm::crossing_t(vol[s], vol[s + o]);
  float c[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
  c[axis] = c[axis] + t;
  verts[3 * (long long)id] = c[0] / (float)R;
  verts[3 * (long long)id + 1] = c[1] / (float)R;
  verts[3 * (long long)id + 2] = c[2] / (float)R;
}
// Per cube code (remesh.h cube_code), and per code the cube's triangle
// and centre counts (nt | nc << 8; 0xFFFF for a code whose segments do not
// close), so the count scans read two bytes per cube.
constexpr uint16_t RM_BAD_CODE = 0xFFFF;
__global__ void k_rm_code_table(uint16_t* __restrict__ tab) {
  const int code = blockIdx.x * blockDim.x + threadIdx.x;
  if (code >= rm::CUBE_CODES) return;
  int next[12], nc = 0;
  if (!rm::cube_segments_code(kMc, code, next)) { tab[code] = RM_BAD_CODE; return; }
  const int nt = rm::cube_tris(kMc, next, nullptr, &nc);
  tab[code] = nt < 0 ? RM_BAD_CODE : (uint16_t)(nt | (nc << 8));
}
__global__ void k_rm_codes(const float* __restrict__ vol, int R, uint16_t* __restrict__ codes) {
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Rc = R - 1;
  if (cell >= (long long)Rc * Rc * Rc) return;
  const int i = (int)(cell % Rc), j = (int)((cell / Rc) % Rc), k = (int)(cell / ((long long)Rc * Rc));
  float val[8];
  for (int c = 0; c < 8; c++)
    val[c] = vol[((long long)(k + ((c >> 2) & 1)) * R + (j + ((c >> 1) & 1))) * R + (i + (c & 1))];
  codes[cell] = (uint16_t)rm::cube_code(kMc, val);
}
// per cube: its triangle count (which 0) or centre count (which 1),
// scanned to the cube's first triangle or centre; a cube whose segments do
// not close counts in err
struct RmCellOp {
  const uint16_t* codes;
  const uint16_t* tab;
  int which;
  int* off;
  int* err;
  struct Rec { int n; };
  __device__ Rec load(long long cell) const {
    const uint16_t t = tab[codes[cell]];
    Rec r;
    if (t == RM_BAD_CODE) {
      if (which == 0) atomicAdd(err, 1);
      r.n = 0;
    } else {
      r.n = which ? (t >> 8) : (t & 0xFF);
    }
    return r;
  }
  __device__ int count(const Rec& r) const { return r.n; }
  __device__ void gather(long long, Rec&) const {}
  __device__ void emit(long long cell, const Rec&, int pos) const { off[cell] = pos; }
};
// the triangles (vertex ids) and centre vertices of every cube
__global__ void k_rm_tris(const uint16_t* __restrict__ codes, const uint16_t* __restrict__ tab, int R,
                          const int* __restrict__ ids, const int* __restrict__ toff, const int* __restrict__ coff,
                          int nedge_verts, float* __restrict__ verts, int* __restrict__ faces) {
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Rc = R - 1;
  if (cell >= (long long)Rc * Rc * Rc) return;
  const int code = codes[cell];
  const uint16_t t = tab[code];
  if (t == RM_BAD_CODE || (t & 0xFF) == 0) return;
  const int i = (int)(cell % Rc), j = (int)((cell / Rc) % Rc), k = (int)(cell / ((long long)Rc * Rc));
  int next[12], nc = 0;
  rm::CubeOut out;
  rm::cube_segments_code(kMc, code, next);
  const int nt = rm::cube_tris(kMc, next, &out, &nc);
  const int c0 = nedge_verts + coff[cell];
  for (int x = 0; x < nc; x++) {   // centre: the mean of its loop's crossings
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int m = 0;
    for (; m < 12 && out.cent[x][m] >= 0; m++) {
      const int id = ids[rm::cube_edge_id(kMc, out.cent[x][m], i, j, k, R)];
      sx = sx + verts[3 * (long long)id];
      sy = sy + verts[3 * (long long)id + 1];
      sz = sz + verts[3 * (long long)id + 2];
    }
    verts[3 * (long long)(c0 + x)] = sx / (float)m;
    verts[3 * (long long)(c0 + x) + 1] = sy / (float)m;
    verts[3 * (long long)(c0 + x) + 2] = sz / (float)m;
  }
  const long long f0 = toff[cell];
  for (int q = 0; q < nt; q++)
    for (int c = 0; c < 3; c++) {
      const int l = out.tris[q][c];
      faces[3 * (f0 + q) + c] = l >= 12 ? c0 + (l - 12) : ids[rm::cube_edge_id(kMc, l, i, j, k, R)];
    }
}

}  // namespace rmk
}  // namespace pamo
