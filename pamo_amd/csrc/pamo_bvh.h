// LBVH over the alive faces and the stackless overlap traversal with the
// reference's narrow phase (simp_cuda/src/bvh/self_intersect.cuh:151-392,
// bvh/tri_tri_3d.cuh:100-297, bvh/math.cuh:43-51).
//
// Build: face AABBs -> scene box -> 30-bit Morton codes of box centres ->
// LDS radix sort of (code, leaf) -> Karras (2012) hierarchy with ties broken
// by leaf index -> bottom-up boxes (agent-scope acq_rel flags) -> ropes.
// Node layout (2 x float4 per node, 32 B):
//   A = (lo.x, lo.y, lo.z, escape)   B = (hi.x, hi.y, hi.z, child)
// child >= 0 is the left child of an internal node, child < 0 is ~face of a
// leaf; escape is the next node in depth-first order after the subtree
// (-1 ends the walk).  Internal nodes are 0..n-2, leaf l is node n-1+l.
// The candidate set (all alive faces whose float AABB overlaps the query
// box, touching included) does not depend on the tree shape.
#pragma once
#include "pamo_kernels.h"
#include "exact_pred.h"

namespace pamo {

// ------------------------------------------------------ narrow phase
__device__ __forceinline__ double dotdf(const double a[3], const float b[3]) {
  return a[0] * (double)b[0] + a[1] * (double)b[1] + a[2] * (double)b[2];
}
__device__ __forceinline__ void crossd(double r[3], const double a[3], const double b[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void plane_of(double pl[4], const float t[3][3]) {
  double ba[3], bc[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ba[k] = (double)t[0][k] - (double)t[1][k];
    bc[k] = (double)t[2][k] - (double)t[1][k];
  }
  crossd(pl, ba, bc);
  pl[3] = -dotdf(pl, t[1]);
}
__device__ __forceinline__ void interp3(float r[3], const float a[3], const float b[3], float t) {
  const float s = 1.0f - t;
  r[0] = s * a[0] + t * b[0];
  r[1] = s * a[1] + t * b[1];
  r[2] = s * a[2] + t * b[2];
}
__device__ __forceinline__ float sel3(int r, float a, float b, float c) { return r == 0 ? a : (r == 1 ? b : c); }
// tri_tri_3d.cuh:131-276, with the first plane and the sides of the second
// triangle passed in (the coplanarity test computes the same values).  The
// reference picks the vertex order (t0, t1, t2) of each triangle at run
// time and indexes with it; here the order is one of three rotations r
// (t0 = r, t1 = r+1, t2 = r+2 mod 3) and the vertices are selected by value,
// so no array is indexed dynamically and nothing goes to scratch memory.
// Every arithmetic expression is the reference's, on the same operands.
__device__ bool isect_tri_tri_ref(const float ta[3][3], const float tb[3][3], const double pa[4], const double sd[3],
                                  float ri1[3], float ri2[3]) {
  float mn[2], mx[2], lo0[2][3], lo1[2][3];
  float side[2][3];
  double pb[4];
#pragma unroll
  for (int k = 0; k < 3; k++) side[1][k] = (float)sd[k];
  if (!side[1][0] && !side[1][1] && !side[1][2]) return false;
  if ((side[1][0] && side[1][1] && side[1][2]) && (side[1][0] < 0.0f) == (side[1][1] < 0.0f) &&
      (side[1][0] < 0.0f) == (side[1][2] < 0.0f))
    return false;
  plane_of(pb, tb);
#pragma unroll
  for (int k = 0; k < 3; k++) side[0][k] = (float)(dotdf(pb, ta[k]) + pb[3]);
  if ((side[0][0] && side[0][1] && side[0][2]) && (side[0][0] < 0.0f) == (side[0][1] < 0.0f) &&
      (side[0][0] < 0.0f) == (side[0][2] < 0.0f))
    return false;
  double dir[3];
  crossd(dir, pa, pb);
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const float(*tri)[3] = i == 0 ? ta : tb;
    const float s0 = side[i][0], s1 = side[i][1], s2 = side[i][2];
    int r;
    if ((s0 && s1) && (s0 < 0.0f) == (s1 < 0.0f)) r = 1;
    else if ((s1 && s2) && (s1 < 0.0f) == (s2 < 0.0f)) r = 2;
    else if (!s1 && (s2 < 0.0f) != (s0 < 0.0f)) r = 1;
    else r = 0;
    float A[3], B[3], C[3];   // tri[t0], tri[t1], tri[t2]
#pragma unroll
    for (int k = 0; k < 3; k++) {
      A[k] = sel3(r, tri[0][k], tri[1][k], tri[2][k]);
      B[k] = sel3(r, tri[1][k], tri[2][k], tri[0][k]);
      C[k] = sel3(r, tri[2][k], tri[0][k], tri[1][k]);
    }
    const float sA = sel3(r, s0, s1, s2), sB = sel3(r, s1, s2, s0), sC = sel3(r, s2, s0, s1);
    double dot_b = dotdf(dir, B);
    if (sB) {
      double dot_a = dotdf(dir, A);
      double dot_c = dotdf(dir, C);
      float fac0 = sB / (sB - sA);
      float fac1 = sB / (sB - sC);
      double off0 = fac0 * (dot_a - dot_b);
      double off1 = fac1 * (dot_c - dot_b);
      if (off0 > off1) {
        double td = off0; off0 = off1; off1 = td;
        float tf = fac0; fac0 = fac1; fac1 = tf;
#pragma unroll
        for (int k = 0; k < 3; k++) { float x = A[k]; A[k] = C[k]; C[k] = x; }
      }
      mn[i] = (float)(dot_b + off0);
      mx[i] = (float)(dot_b + off1);
      interp3(lo0[i], B, A, fac0);
      interp3(lo1[i], B, C, fac1);
    } else {
      mn[i] = mx[i] = (float)dot_b;
#pragma unroll
      for (int k = 0; k < 3; k++) lo0[i][k] = lo1[i][k] = B[k];
    }
  }
  if ((mx[0] > mn[1]) && (mn[0] < mx[1])) {
    // the reference's choice of segment ends, by value
    const bool c1 = mn[0] >= mn[1], c2 = mx[0] <= mx[1];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      ri1[k] = c1 ? lo0[0][k] : (c2 ? lo1[0][k] : lo0[1][k]);
      ri2[k] = c1 ? (c2 ? lo1[0][k] : lo1[1][k]) : (c2 ? lo0[1][k] : lo1[1][k]);
    }
    return true;
  }
  return false;
}

// One ordered pair (current c, query q): self_intersect.cuh:233-378.
// cur/qry are the raw (untranslated) vertex positions; nc the number of
// equal vertex-index pairs.  Callers skip nc == 2 (both branches of the
// reference return false for it, :285-286 and :338-339).
__device__ bool narrow_hit(const float cur[3][3], const float qry[3][3], int nc) {
  float ta[3][3], tb[3][3];
  int mean[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float m = (cur[0][k] + cur[1][k] + cur[2][k] + qry[0][k] + qry[1][k] + qry[2][k]) / 6.0f;
    mean[k] = (int)floorf(m);
  }
  float scale = 0.0f;
#pragma unroll
  for (int v = 0; v < 3; v++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      ta[v][k] = cur[v][k] - (float)mean[k];
      tb[v][k] = qry[v][k] - (float)mean[k];
    }
  // is_coplanar (tri_tri_3d.cuh:100-129): the same plane and sides feed the
  // interval test
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) scale = fmaxf(scale, fabsf(ta[i][j]));
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) scale = fmaxf(scale, fabsf(tb[i][j]));
  double pa[4], sd[3];
  plane_of(pa, ta);
#pragma unroll
  for (int k = 0; k < 3; k++) sd[k] = dotdf(pa, tb[k]) + pa[3];
  const double eps = (double)scale * 1e-6;
  if (fabs(sd[0]) <= eps && fabs(sd[1]) <= eps && fabs(sd[2]) <= eps) return false;
  if (nc == 2) return false;
  {
    const float e6 = 1e-6f;
    int sh = 0;
#pragma unroll
    for (int v = 0; v < 3; v++) {
      bool any = false;
#pragma unroll
      for (int w = 0; w < 3; w++)
        any = any || (fabsf(ta[v][0] - tb[w][0]) < e6 && fabsf(ta[v][1] - tb[w][1]) < e6 &&
                      fabsf(ta[v][2] - tb[w][2]) < e6);
      sh += any;
    }
    if (sh >= 2) return false;
  }
  float r1[3], r2[3];
  if (!isect_tri_tri_ref(ta, tb, pa, sd, r1, r2)) return false;
  float dx = fabsf(r2[0] - r1[0]), dy = fabsf(r2[1] - r1[1]), dz = fabsf(r2[2] - r1[2]);
  float ld = -INFINITY;
  if (dx > ld) ld = dx;
  if (dy > ld) ld = dy;
  if (dz > ld) ld = dz;
  if (ld < 1e-3f && nc == 1) return false;
  return true;
}
__device__ __forceinline__ int shared_indices(I3 a, I3 b) {
  int av[3] = {a.i, a.j, a.k}, bv[3] = {b.i, b.j, b.k};
  int nc = 0;
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) nc += av[j] == bv[k];
  return nc;
}
// both orientations of a candidate pair (query f, other g)
__device__ __forceinline__ bool pair_hits(const float pf[3][3], const float pg[3][3], I3 tf, I3 tg) {
  int nc = shared_indices(tg, tf);
  if (nc == 2) return false;
  return narrow_hit(pg, pf, nc) || narrow_hit(pf, pg, nc);
}
// The strict guard (not in the reference; off by default): a pair also
// counts when the exact closed triangle test of exact_pred.h says the faces
// intersect, which decides the pairs the reference's narrow phase skips
// (coplanar, shared-edge coordinates, short shared-vertex contacts) and
// those its float interval test misses.
__device__ __attribute__((noinline)) bool exact_pair(const float pf[3][3], const float pg[3][3], I3 tf, I3 tg) {
  xp::P3 A[3], B[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    A[k] = {(double)pf[k][0], (double)pf[k][1], (double)pf[k][2]};
    B[k] = {(double)pg[k][0], (double)pg[k][1], (double)pg[k][2]};
  }
  const int ia[3] = {tf.i, tf.j, tf.k}, ib[3] = {tg.i, tg.j, tg.k};
  return xp::tri_pair(A, ia, B, ib);
}
// STRICT is a template parameter: the default kernels never reference the
// exact test, which would otherwise set their register budget to its own
template <int STRICT>
__device__ __forceinline__ bool pair_hits_g(const float pf[3][3], const float pg[3][3], I3 tf, I3 tg) {
  if (pair_hits(pf, pg, tf, tg)) return true;
  if constexpr (STRICT != 0) return exact_pair(pf, pg, tf, tg);
  return false;
}

// ------------------------------------------------------------- LBVH
// leaf boxes for the alive-face list; scene box by block reduction into
// 6 ordered-int slots (min via atomicMin on the float's ordered encoding)
__device__ __forceinline__ int f2ord(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void k_scene_init(int* scene) {
  if (threadIdx.x < 3) scene[threadIdx.x] = 0x7FFFFFFF;
  else if (threadIdx.x < 6) scene[threadIdx.x] = (int)0x80000000;
}

// Leaf boxes of the LBVH builds are interleaved records: the low corner of
// leaf l at lbox[2 l] (its face in .w), the high corner at lbox[2 l + 1];
// callers pass lbox_lo = lbox and lbox_hi = lbox + 1.  The node boxes gather
// the leaves through the Morton order, one 32-byte record per leaf.
constexpr int LBOX_STRIDE = 2;
__global__ void __launch_bounds__(BLOCK) k_leaf_boxes(const int* __restrict__ leaf_face, const int* __restrict__ n_ptr,
                                                      const float* __restrict__ P, const int* __restrict__ T,
                                                      float4* __restrict__ lbox_lo, float4* __restrict__ lbox_hi,
                                                      int* scene) {
  __shared__ float red[6][BLOCK / 64];
  int n = *n_ptr;
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (l < n) {
    const int f = leaf_face[l];
    I3 t = ldt(T, f);
    V3 a = ldp(P, t.i), b = ldp(P, t.j), c = ldp(P, t.k);
    lo[0] = fminf(a.x, fminf(b.x, c.x)); lo[1] = fminf(a.y, fminf(b.y, c.y)); lo[2] = fminf(a.z, fminf(b.z, c.z));
    hi[0] = fmaxf(a.x, fmaxf(b.x, c.x)); hi[1] = fmaxf(a.y, fmaxf(b.y, c.y)); hi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
    lbox_lo[LBOX_STRIDE * l] = make_float4(lo[0], lo[1], lo[2], __int_as_float(f));   // .w: the leaf's face
    lbox_hi[LBOX_STRIDE * l] = make_float4(hi[0], hi[1], hi[2], 0.0f);
  }
  float v[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    float x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      float y = __shfl_xor(x, o, 64);
      x = k < 3 ? fminf(x, y) : fmaxf(x, y);
    }
    v[k] = x;
  }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; k++) red[k][threadIdx.x >> 6] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 6; k++) {
      float x = red[k][0];
      for (int w = 1; w < BLOCK / 64; w++) x = k < 3 ? fminf(x, red[k][w]) : fmaxf(x, red[k][w]);
      if (x == x && x != INFINITY && x != -INFINITY) {
        if (k < 3) atomicMin(&scene[k], f2ord(x));
        else atomicMax(&scene[k], f2ord(x));
      }
    }
  }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
__global__ void k_morton(const float4* __restrict__ lbox_lo, const float4* __restrict__ lbox_hi,
                         const int* __restrict__ n_ptr, const int* __restrict__ scene, uint32_t* __restrict__ keys,
                         int* __restrict__ vals) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int n = *n_ptr;
  if (l >= n) return;
  float slo[3] = {ord2f(scene[0]), ord2f(scene[1]), ord2f(scene[2])};
  float shi[3] = {ord2f(scene[3]), ord2f(scene[4]), ord2f(scene[5])};
  float4 lo = lbox_lo[LBOX_STRIDE * l], hi = lbox_hi[LBOX_STRIDE * l];
  float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float ext = shi[k] - slo[k];
    float t = ext > 0.0f ? (c[k] - slo[k]) / ext : 0.0f;
    t = fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
    q[k] = (uint32_t)t;
  }
  keys[l] = expand10(q[0]) * 4 + expand10(q[1]) * 2 + expand10(q[2]);
  vals[l] = l;
}

// -------------------------------------------------------- radix sort
// Stable LSD radix sort of (uint32 key, int value), 8-bit digits.
// Pass = per-tile digit histograms -> global scan (digit-major) -> scatter
// with in-tile stable ranks from an LDS 1-bit split sort of the tile.
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = BLOCK * RS_ITEMS;

__global__ void __launch_bounds__(BLOCK) k_rs_hist(const uint32_t* __restrict__ keys, const int* __restrict__ n_ptr,
                                                   int shift, int ntiles, int* __restrict__ hist) {
  __shared__ int h[256];
  int n = *n_ptr;
  h[threadIdx.x] = 0;
  __syncthreads();
  long long base = (long long)blockIdx.x * RS_TILE;
  for (int i = 0; i < RS_ITEMS; i++) {
    long long idx = base + (long long)i * BLOCK + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(BLOCK) k_rs_scatter(const uint32_t* __restrict__ kin, const int* __restrict__ vin,
                                                      const int* __restrict__ n_ptr, int shift, int ntiles,
                                                      const int* __restrict__ hoff, uint32_t* __restrict__ kout,
                                                      int* __restrict__ vout) {
  __shared__ uint32_t sk[RS_TILE];
  __shared__ int sv[RS_TILE];
  __shared__ int lds[BLOCK / 64];
  __shared__ int dstart[256];
  __shared__ int goff[256];
  int n = *n_ptr;
  long long base = (long long)blockIdx.x * RS_TILE;
  if (base >= n) return;
  int cnt = (int)min((long long)RS_TILE, (long long)n - base);
  // blocked load: thread t owns tile items t*ITEMS .. t*ITEMS+ITEMS-1
  uint32_t k[RS_ITEMS];
  int v[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int li = threadIdx.x * RS_ITEMS + i;
    if (li < cnt) { k[i] = kin[base + li]; v[i] = vin[base + li]; }
    else { k[i] = 0xFFFFFFFFu; v[i] = -1; }   // pads sort to the end
  }
  goff[threadIdx.x] = hoff[(long long)threadIdx.x * ntiles + blockIdx.x];
  // 8 stable 1-bit splits of the tile on the digit bits
  for (int bit = 0; bit < 8; bit++) {
    int ones = 0;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) ones += (int)((k[i] >> (shift + bit)) & 1u);
    int zeros_here = RS_ITEMS - ones;
    int total_zeros;
    int zbefore = block_exclusive_scan(zeros_here, lds, total_zeros);
    int obefore = threadIdx.x * RS_ITEMS - zbefore;
    int zc = 0, oc = 0;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
      int b = (int)((k[i] >> (shift + bit)) & 1u);
      int dst = b ? (total_zeros + obefore + oc++) : (zbefore + zc++);
      sk[dst] = k[i];
      sv[dst] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) { k[i] = sk[threadIdx.x * RS_ITEMS + i]; v[i] = sv[threadIdx.x * RS_ITEMS + i]; }
    __syncthreads();
  }
  // digit start positions inside the sorted tile
  dstart[threadIdx.x] = RS_TILE;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int li = threadIdx.x * RS_ITEMS + i;
    if (li < cnt) {
      int d = (int)((k[i] >> shift) & 255u);
      int prev = li == 0 ? -1 : (int)((sk[li - 1] >> shift) & 255u);
      if (li == 0 || prev != d) dstart[d] = li;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int li = threadIdx.x * RS_ITEMS + i;
    if (li < cnt) {
      int d = (int)((k[i] >> shift) & 255u);
      long long o = (long long)goff[d] + (li - dstart[d]);
      kout[o] = k[i];
      vout[o] = v[i];
    }
  }
}

// --------------------------------------------------------------- Karras
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ key, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  uint32_t a = key[i], b = key[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}
// internal node i in [0, n-2]; writes children, parents and ranges
__global__ void k_karras(const uint32_t* __restrict__ key, const int* __restrict__ n_ptr, int* __restrict__ left,
                         int* __restrict__ right, int* __restrict__ parent, int* __restrict__ rfirst,
                         int* __restrict__ rlast) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = *n_ptr;
  if (i >= n - 1) return;
  int d = (kdelta(key, n, i, i + 1) - kdelta(key, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = kdelta(key, n, i, i - d);
  int lmax = 2;
  while (kdelta(key, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(key, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = kdelta(key, n, i, j);
  int s = 0;
  for (int div = 2;; div *= 2) {
    int t = (l + div - 1) / div;
    if (kdelta(key, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  int g = i + s * d + (d < 0 ? -1 : 0);
  int lo = min(i, j), hi = max(i, j);
  int lc = (lo == g) ? (n - 1 + g) : g;
  int rc = (hi == g + 1) ? (n - 1 + g + 1) : (g + 1);
  left[i] = lc;
  right[i] = rc;
  parent[lc] = i;
  parent[rc] = i;
  rfirst[i] = lo;
  rlast[i] = hi;
}

// Bottom-up node boxes in two phases.  Phase 1: one block per BLOCK sorted
// leaves; nodes whose leaf range lies inside the block are merged through
// LDS flags (workgroup scope: the second child to arrive merges).  Phase 2:
// the few nodes whose range spans blocks, from their in-block children,
// through agent-scope acq_rel flags in global memory.
__device__ __forceinline__ long long seg_base(int j, long long cap) {
  long long b = 0;
  for (int i = 1; i < j; i++) b += (cap + (1LL << i) - 1) >> i;
  return b;
}
inline long long seg_entries(long long cap) {   // host: buffer entries for all levels
  long long b = 0;
  for (int i = 1; (1LL << (i - 1)) < cap; i++) b += (cap + (1LL << i) - 1) >> i;
  return b + 1;
}
// block reduction of BLOCK boxes (lo/hi in LDS, NaN padding) into levels
// j0+1..j0+8; entry index of the block's first element at level j0 is e0
__device__ __forceinline__ void seg_reduce_block(float (*rlo)[3], float (*rhi)[3], long long e0, int j0, int n,
                                                 long long cap, float4* __restrict__ seg_lo,
                                                 float4* __restrict__ seg_hi) {
  const int t = threadIdx.x;
  long long base = seg_base(j0 + 1, cap);
  for (int d = 1; (BLOCK >> d) >= 1; d++) {
    const int j = j0 + d;
    const int cnt = BLOCK >> d;
    float lo[3], hi[3];
    const bool act = t < cnt;
    if (act) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = fminf(rlo[2 * t][k], rlo[2 * t + 1][k]);
        hi[k] = fmaxf(rhi[2 * t][k], rhi[2 * t + 1][k]);
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int k = 0; k < 3; k++) { rlo[t][k] = lo[k]; rhi[t][k] = hi[k]; }
      long long idx = (e0 >> d) + t;
      long long size = ((long long)n + (1LL << j) - 1) >> j;
      if (idx < size && (1LL << (j - 1)) < n) {   // levels a query can use (2^(j-1) < n <= cap)
        seg_lo[base + idx] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        seg_hi[base + idx] = make_float4(hi[0], hi[1], hi[2], 0.0f);
      }
    }
    __syncthreads();
    base += (cap + (1LL << j) - 1) >> j;
  }
}

__device__ __forceinline__ int escape_after(int last, int n, const int* __restrict__ rfirst) {
  // next node in depth-first order after a subtree ending at sorted leaf last
  if (last == n - 1) return -1;
  return (last + 1 <= n - 2 && rfirst[last + 1] == last + 1) ? (last + 1) : (n - 1 + last + 1);
}
__device__ __forceinline__ void node_range(int x, int n, const int* __restrict__ rfirst,
                                           const int* __restrict__ rlast, int& a, int& b) {
  if (x >= n - 1) { a = b = x - (n - 1); }
  else { a = rfirst[x]; b = rlast[x]; }
}

__global__ void __launch_bounds__(BLOCK) k_bottom_up_local(const int* __restrict__ n_ptr,
                                                           const float4* __restrict__ lbox_lo,
                                                           const float4* __restrict__ lbox_hi,
                                                           const int* __restrict__ sorted_leaf,
                                                           const int* __restrict__ parent, const int* __restrict__ left,
                                                           const int* __restrict__ right, const int* __restrict__ rfirst,
                                                           const int* __restrict__ rlast, float4* nodeA, float4* nodeB,
                                                           long long cap, float4* __restrict__ seg_lo,
                                                           float4* __restrict__ seg_hi) {
  __shared__ int flag[BLOCK];
  __shared__ float blo[2 * BLOCK][3], bhi[2 * BLOCK][3];   // [0,B): leaves, [B,2B): internal nodes
  __shared__ float rlo[BLOCK][3], rhi[BLOCK][3];           // segment-tree reduction
  const int n = *n_ptr;
  const int b0 = blockIdx.x * BLOCK;
  if (b0 >= n) return;
  const int b1 = min(b0 + BLOCK, n) - 1;
  flag[threadIdx.x] = 0;
  const int l = b0 + threadIdx.x;
  if (l <= b1) {
    int src = sorted_leaf[l];
    float4 lo = lbox_lo[LBOX_STRIDE * src], hi = lbox_hi[LBOX_STRIDE * src];
    blo[threadIdx.x][0] = lo.x; blo[threadIdx.x][1] = lo.y; blo[threadIdx.x][2] = lo.z;
    bhi[threadIdx.x][0] = hi.x; bhi[threadIdx.x][1] = hi.y; bhi[threadIdx.x][2] = hi.z;
    nodeA[n - 1 + l] = make_float4(lo.x, lo.y, lo.z, __int_as_float(escape_after(l, n, rfirst)));
    // (the leaf's face rides in lo.w: one gathered record fewer per leaf)
    nodeB[n - 1 + l] = make_float4(hi.x, hi.y, hi.z, __int_as_float(~__float_as_int(lo.w)));
    rlo[threadIdx.x][0] = lo.x; rlo[threadIdx.x][1] = lo.y; rlo[threadIdx.x][2] = lo.z;
    rhi[threadIdx.x][0] = hi.x; rhi[threadIdx.x][1] = hi.y; rhi[threadIdx.x][2] = hi.z;
  } else {
    const float qn = __int_as_float(0x7FC00000);
#pragma unroll
    for (int k = 0; k < 3; k++) rlo[threadIdx.x][k] = rhi[threadIdx.x][k] = qn;
  }
  __syncthreads();
  if (n > BLOCK) seg_reduce_block(rlo, rhi, b0, 0, n, cap, seg_lo, seg_hi);   // block-uniform
  if (l > b1 || n == 1) return;
  int node = n - 1 + l;
  for (int depth = 0; depth < n; depth++) {
    int p = parent[node];
    if (p < 0) return;
    int pa = rfirst[p], pb = rlast[p];
    if (pa < b0 || pb > b1) return;   // spans blocks: phase 2
    int old = __hip_atomic_fetch_add(&flag[p - b0], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0) return;
    int lc = left[p], rc = right[p];
    int li = lc >= n - 1 ? (lc - (n - 1) - b0) : (BLOCK + lc - b0);
    int ri = rc >= n - 1 ? (rc - (n - 1) - b0) : (BLOCK + rc - b0);
    float lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = fminf(blo[li][k], blo[ri][k]);
      hi[k] = fmaxf(bhi[li][k], bhi[ri][k]);
    }
    int pi = BLOCK + p - b0;
#pragma unroll
    for (int k = 0; k < 3; k++) { blo[pi][k] = lo[k]; bhi[pi][k] = hi[k]; }
    nodeA[p] = make_float4(lo[0], lo[1], lo[2], __int_as_float(escape_after(pb, n, rfirst)));
    nodeB[p] = make_float4(hi[0], hi[1], hi[2], __int_as_float(lc));
    node = p;
  }
}

// Nodes whose leaf range spans blocks.  A node's box is the union of the
// leaf boxes in its range (fminf/fmaxf ignore NaN and are otherwise exact,
// so any grouping gives the box of the pairwise bottom-up merge), taken here
// from an aligned segment tree over the sorted leaves: level j >= 1, entry i
// is the union of leaves [i 2^j, (i+1) 2^j) clipped to n, stored at
// seg_base(j) + i with bases laid out for a capacity of cap leaves.
// Levels 1..8 come from k_bottom_up_local (one block = 256 aligned leaves),
// higher ones from k_seg_upper, 8 levels per launch.  No node waits for
// another, so no cross-workgroup flags or cache maintenance are needed.
__global__ void __launch_bounds__(BLOCK) k_seg_upper(const int* __restrict__ n_ptr, long long cap, int j0,
                                                     float4* __restrict__ seg_lo, float4* __restrict__ seg_hi) {
  __shared__ float rlo[BLOCK][3], rhi[BLOCK][3];
  const int n = *n_ptr;
  const long long size0 = ((long long)n + (1LL << j0) - 1) >> j0;
  const long long e0 = (long long)blockIdx.x * BLOCK;
  if (size0 <= 1 || e0 >= size0) return;   // block-uniform
  const long long e = e0 + threadIdx.x;
  const float qn = __int_as_float(0x7FC00000);
  const long long b0 = seg_base(j0, cap);
#pragma unroll
  for (int k = 0; k < 3; k++) rlo[threadIdx.x][k] = rhi[threadIdx.x][k] = qn;
  if (e < size0) {
    float4 lo = seg_lo[b0 + e], hi = seg_hi[b0 + e];
    rlo[threadIdx.x][0] = lo.x; rlo[threadIdx.x][1] = lo.y; rlo[threadIdx.x][2] = lo.z;
    rhi[threadIdx.x][0] = hi.x; rhi[threadIdx.x][1] = hi.y; rhi[threadIdx.x][2] = hi.z;
  }
  __syncthreads();
  seg_reduce_block(rlo, rhi, e0, j0, n, cap, seg_lo, seg_hi);
}

__global__ void __launch_bounds__(BLOCK) k_cross_nodes(const int* __restrict__ n_ptr, long long cap,
                                                       const int* __restrict__ left, const int* __restrict__ rfirst,
                                                       const int* __restrict__ rlast,
                                                       const float4* __restrict__ seg_lo,
                                                       const float4* __restrict__ seg_hi, float4* nodeA,
                                                       float4* nodeB) {
  const int n = *n_ptr;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < 2 || i >= n - 1) return;
  const int a = rfirst[i], b = rlast[i];
  if (a / BLOCK == b / BLOCK) return;   // merged in k_bottom_up_local
  const float qn = __int_as_float(0x7FC00000);
  float lo[3] = {qn, qn, qn}, hi[3] = {qn, qn, qn};
  long long l = a, h = (long long)b + 1;
  long long base = 0;   // base of level j (j >= 1)
  for (int j = 0; l < h; j++) {
    if (l & 1) {
      float4 x, y;
      if (j == 0) { x = nodeA[n - 1 + l]; y = nodeB[n - 1 + l]; }
      else { x = seg_lo[base + l]; y = seg_hi[base + l]; }
      lo[0] = fminf(lo[0], x.x); lo[1] = fminf(lo[1], x.y); lo[2] = fminf(lo[2], x.z);
      hi[0] = fmaxf(hi[0], y.x); hi[1] = fmaxf(hi[1], y.y); hi[2] = fmaxf(hi[2], y.z);
      l++;
    }
    if (h & 1) {
      h--;
      float4 x, y;
      if (j == 0) { x = nodeA[n - 1 + h]; y = nodeB[n - 1 + h]; }
      else { x = seg_lo[base + h]; y = seg_hi[base + h]; }
      lo[0] = fminf(lo[0], x.x); lo[1] = fminf(lo[1], x.y); lo[2] = fminf(lo[2], x.z);
      hi[0] = fmaxf(hi[0], y.x); hi[1] = fmaxf(hi[1], y.y); hi[2] = fmaxf(hi[2], y.z);
    }
    l >>= 1;
    h >>= 1;
    if (j >= 1) base += (cap + (1LL << j) - 1) >> j;
  }
  nodeA[i] = make_float4(lo[0], lo[1], lo[2], __int_as_float(escape_after(b, n, rfirst)));
  nodeB[i] = make_float4(hi[0], hi[1], hi[2], __int_as_float(left[i]));
}

// ------------------------------------------------------------ traversal
// One thread per query face: walk the ropes, run both orientations of the
// narrow phase on every overlapping alive face, S[f] = 1 on the first hit.
// A walk visits each node at most once; the step cap only guards against a
// malformed tree (counted in err, never expected).
__global__ void __launch_bounds__(BLOCK) k_traverse(const int* __restrict__ qlist, const int* __restrict__ nq_ptr,
                                                    const int* __restrict__ n_ptr, const float4* __restrict__ nodeA,
                                                    const float4* __restrict__ nodeB, const float* __restrict__ P,
                                                    const int* __restrict__ T, uint8_t* __restrict__ S,
                                                    unsigned long long* __restrict__ stats, int* err) {
  int qi = blockIdx.x * blockDim.x + threadIdx.x;
  int nq = *nq_ptr;
  unsigned cand = 0;
  bool hit = false;
  if (qi < nq) {
    const long long cap = 2LL * (2LL * (*n_ptr) - 1) + 8;
    int f = qlist[qi];
    I3 tf = ldt(T, f);
    float pf[3][3];
    {
      V3 a = ldp(P, tf.i), b = ldp(P, tf.j), c = ldp(P, tf.k);
      pf[0][0] = a.x; pf[0][1] = a.y; pf[0][2] = a.z;
      pf[1][0] = b.x; pf[1][1] = b.y; pf[1][2] = b.z;
      pf[2][0] = c.x; pf[2][1] = c.y; pf[2][2] = c.z;
    }
    float qlo[3], qhi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      qlo[k] = fminf(pf[0][k], fminf(pf[1][k], pf[2][k]));
      qhi[k] = fmaxf(pf[0][k], fmaxf(pf[1][k], pf[2][k]));
    }
    int node = 0;
    long long steps = 0;
    while (node != -1) {
      if (++steps > cap) { atomicAdd(err, 1); break; }
      float4 A = nodeA[node], B = nodeB[node];
      bool ov = !(qhi[0] < A.x || B.x < qlo[0]) && !(qhi[1] < A.y || B.y < qlo[1]) && !(qhi[2] < A.z || B.z < qlo[2]);
      int esc = __float_as_int(A.w);
      if (ov) {
        int c = __float_as_int(B.w);
        if (c >= 0) { node = c; continue; }
        int g = ~c;
        if (g != f) {
          cand++;
          I3 tg = ldt(T, g);
          float pg[3][3];
          V3 a = ldp(P, tg.i), b = ldp(P, tg.j), cc = ldp(P, tg.k);
          pg[0][0] = a.x; pg[0][1] = a.y; pg[0][2] = a.z;
          pg[1][0] = b.x; pg[1][1] = b.y; pg[1][2] = b.z;
          pg[2][0] = cc.x; pg[2][1] = cc.y; pg[2][2] = cc.z;
          if (pair_hits(pf, pg, tf, tg)) { hit = true; break; }
        }
      }
      node = esc;
    }
    S[f] = hit ? 1 : 0;
  }
  if (stats) {
    unsigned long long c64 = cand, h64 = hit ? 1ull : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      c64 += __shfl_xor(c64, o, 64);
      h64 += __shfl_xor(h64, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && (c64 || h64)) {
      atomicAdd(&stats[0], c64);
      atomicAdd(&stats[1], h64);
    }
  }
}

// Debug: every ordered candidate pair (q, c) and hit flag, for a query list
// (counts first, then a second pass writes at offsets).
__global__ void k_traverse_list(const int* __restrict__ qlist, const int* __restrict__ nq_ptr,
                                const float4* __restrict__ nodeA, const float4* __restrict__ nodeB,
                                const float* __restrict__ P, const int* __restrict__ T, int* __restrict__ counts,
                                const int* __restrict__ offsets, int* __restrict__ pairs) {
  int qi = blockIdx.x * blockDim.x + threadIdx.x;
  int nq = *nq_ptr;
  if (qi >= nq) return;
  int f = qlist[qi];
  I3 tf = ldt(T, f);
  float pf[3][3];
  {
    V3 a = ldp(P, tf.i), b = ldp(P, tf.j), c = ldp(P, tf.k);
    pf[0][0] = a.x; pf[0][1] = a.y; pf[0][2] = a.z;
    pf[1][0] = b.x; pf[1][1] = b.y; pf[1][2] = b.z;
    pf[2][0] = c.x; pf[2][1] = c.y; pf[2][2] = c.z;
  }
  float qlo[3], qhi[3];
  for (int k = 0; k < 3; k++) {
    qlo[k] = fminf(pf[0][k], fminf(pf[1][k], pf[2][k]));
    qhi[k] = fmaxf(pf[0][k], fmaxf(pf[1][k], pf[2][k]));
  }
  int node = 0, cnt = 0;
  int o = offsets ? offsets[qi] : 0;
  while (node != -1) {
    float4 A = nodeA[node], B = nodeB[node];
    bool ov = !(qhi[0] < A.x || B.x < qlo[0]) && !(qhi[1] < A.y || B.y < qlo[1]) && !(qhi[2] < A.z || B.z < qlo[2]);
    int esc = __float_as_int(A.w);
    if (ov) {
      int c = __float_as_int(B.w);
      if (c >= 0) { node = c; continue; }
      int g = ~c;
      if (g != f) {
        if (pairs) {
          I3 tg = ldt(T, g);
          float pg[3][3];
          V3 a = ldp(P, tg.i), b = ldp(P, tg.j), cc = ldp(P, tg.k);
          pg[0][0] = a.x; pg[0][1] = a.y; pg[0][2] = a.z;
          pg[1][0] = b.x; pg[1][1] = b.y; pg[1][2] = b.z;
          pg[2][0] = cc.x; pg[2][1] = cc.y; pg[2][2] = cc.z;
          pairs[3 * (o + cnt)] = f;
          pairs[3 * (o + cnt) + 1] = g;
          pairs[3 * (o + cnt) + 2] = (shared_indices(tg, tf) != 2 && narrow_hit(pg, pf, shared_indices(tg, tf))) ? 1 : 0;
        }
        cnt++;
      }
    }
    node = esc;
  }
  if (counts) counts[qi] = cnt;
}

}  // namespace pamo
