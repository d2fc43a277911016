// gfx950 kernels of the stage-2 round (QEM edge collapse + LBVH guard).
//
// Float semantics follow the reference exactly (simp_cuda/src/cusimp_free.cu,
// bvh/tri_tri_3d.cuh, bvh/self_intersect.cuh): the library is compiled with
// -ffp-contract=off, IEEE division and sqrt, no denormal flushing, so every
// value below is bit-identical to the reference's expression order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pamo {

typedef unsigned long long u64;

constexpr int BLOCK = 256;
constexpr float COST_RANGE = 10.0f;          // cusimp_free.cu:15
constexpr float U32F = 4294967296.0f;        // (float)UINT32_MAX

struct V3 { float x, y, z; };
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float vnorm(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ V3 vdivs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// cusimp_free.cu:29-34 (NaN maps to the upper bound)
__device__ __forceinline__ float clamp_ref(float x, float lo, float hi) {
  float m = lo > x ? lo : x;
  return m < hi ? m : hi;
}
__device__ __forceinline__ uint32_t u32sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= U32F) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

__device__ __forceinline__ V3 ldp(const float* __restrict__ P, int v) {
  return {P[3 * v], P[3 * v + 1], P[3 * v + 2]};
}
__device__ __forceinline__ void stp(float* P, int v, V3 p) {
  P[3 * v] = p.x; P[3 * v + 1] = p.y; P[3 * v + 2] = p.z;
}
struct I3 { int i, j, k; };
__device__ __forceinline__ I3 ldt(const int* __restrict__ T, int f) {
  return {T[3 * f], T[3 * f + 1], T[3 * f + 2]};
}
__device__ __forceinline__ bool dead3(I3 t) { return t.i == -1 && t.j == -1 && t.k == -1; }

// Work enqueued ahead of a decision the device makes (a round queued behind
// one whose result the host has not read) skips itself, writing nothing,
// while *gate is non-zero; a null gate never skips.  The gate is written by
// an earlier launch on the same stream, so it is uniform over the grid.
__device__ __forceinline__ bool gated(const int* gate) { return gate && *gate; }

// ------------------------------------------------------------ generic
__global__ void k_fill_i32(int* p, int v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(u64* p, u64 v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, uint8_t v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------- block scan
// Exclusive scan of one int per thread across a block of BLOCK threads
// (wave64 shuffles, then one LDS exchange of the four wave totals).
__device__ __forceinline__ int block_exclusive_scan(int x, int* lds, int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) lds[wid] = v;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; w++) {
    int s = lds[w];
    if (w < wid) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + v - x;
}

// the same for a block of NT threads (unsigned values)
constexpr int BLOCK1K = 1024;
template <int NT>
__device__ __forceinline__ unsigned block_exclusive_scan_nt(unsigned x, unsigned* lds, unsigned& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) lds[wid] = v;
  __syncthreads();
  unsigned off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) {
    unsigned s = lds[w];
    if (w < wid) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + v - x;
}
__device__ __forceinline__ unsigned block_exclusive_scan_1024(unsigned x, unsigned* lds, unsigned& total) {
  return block_exclusive_scan_nt<BLOCK1K>(x, lds, total);
}

// ------------------------------------------------ one-pass exclusive scan
// Decoupled look-back: tile t (one block, NT x ITEMS ints) scans itself,
// publishes its aggregate, adds the aggregates/prefixes of earlier tiles
// and publishes its inclusive prefix.  A tile's status and value travel in
// one 64-bit word written and read with relaxed device-scope atomics, so no
// ordering between separate locations (and no L2 write-back or invalidate)
// is needed.  Words carry the call's epoch, so stale words from earlier
// calls read as "not yet published" and the array never needs clearing
// (except when the 22-bit epoch wraps, which the host handles).  Blocks are
// dispatched in index order, so every tile a block waits on is running or
// done; the spin is capped anyway and reports through err.  Inputs are
// non-negative (counts and flags); totals stay below 2^40.
constexpr unsigned long long SCAN_ST_AGG = 1ull, SCAN_ST_PREFIX = 2ull;
// Tile shapes (threads x items per thread).  A tile is one workgroup on one
// CU, so short inputs take small tiles to spread over more CUs; long ones
// take large tiles for fewer look-back steps.  Measured on the MI355X (edge
// list of 16k faces: 7.9 us with 512x4 tiles, 18.8 us with 1024x8; at 2M
// faces both 77 us; plain array scans gain only below 64k items).
constexpr int SCAN1_ITEMS = 8, SCAN1_TILE = BLOCK1K * SCAN1_ITEMS;   // 8192 ints per tile
constexpr int SCANS_NT = 512, SCANS_ITEMS = 4, SCANS_TILE = SCANS_NT * SCANS_ITEMS;   // 2048
constexpr long long SCAN_SMALL_MAX = 65536;        // small tiles up to this length ...
constexpr long long SCAN_SMALL_MAX_HEAVY = 1LL << 20;   // ... or this, for ops with gathers and stores per item
__device__ __forceinline__ unsigned long long scan_word(unsigned epoch, unsigned long long st, long long v) {
  return ((unsigned long long)epoch << 42) | (st << 40) | ((unsigned long long)v & ((1ull << 40) - 1));
}
// The scan takes its input from the op and hands each exclusive prefix to
// op.emit, so a producer (counts from faces) and a consumer (a scatter to
// the prefix position) can share the scan's launch.  Per item the op
// provides
//   Rec load(i)             the item's record (its loads)
//   int count(rec)          the value scanned
//   void gather(i, rec)     further loads and atomics of the emit
//   void emit(i, rec, pos)  the stores
// The kernel runs each phase for all of a thread's items before the next,
// so no item's loads wait behind another item's stores (which the compiler
// must assume may alias them); the gathers are issued before the
// look-back and overlap it.
// A non-zero *gate skips the scan and writes nothing (gated()), so work can
// be enqueued behind a decision the device makes.
// A scan recorded into a graph that is replayed takes its epoch from the
// device: the call's index within the graph plus *epoch_base, which moves on
// by SCAN_EPOCH_STEP every replay (multiples of the step keep the sum
// non-zero within the 22 bits).
constexpr unsigned SCAN_EPOCH_MASK = (1u << 22) - 1, SCAN_EPOCH_STEP = 64;
template <class Op, int NT = BLOCK1K, int ITEMS = SCAN1_ITEMS>
__global__ void __launch_bounds__(NT) k_scan_op(Op op, long long n, const int* __restrict__ d_n, int add,
                                                unsigned long long* desc, unsigned epoch,
                                                const unsigned* __restrict__ epoch_base, int* __restrict__ total,
                                                int* err, const int* __restrict__ gate) {
  __shared__ unsigned lds[NT / 64];
  constexpr int TILE = NT * ITEMS;
  __shared__ int tile[TILE];
  __shared__ long long excl_s;
  if (gated(gate)) return;
  if (epoch_base) epoch = (epoch + *epoch_base) & SCAN_EPOCH_MASK;
  const long long live = d_n ? min(n, (long long)*d_n + add) : n;
  const long long t = blockIdx.x;
  const long long base = t * TILE;
  if (live <= 0) {
    if (t == 0 && threadIdx.x == 0 && total) *total = 0;
    return;
  }
  if (base >= live) return;   // tiles past the live length contribute nothing
  typename Op::Rec rec[ITEMS];
  // unrolled and branch-free (items past the end load the last live item
  // and count 0), so the items' loads are in flight together
#pragma unroll
  for (int i = 0; i < ITEMS; i++) {
    long long idx = base + (long long)i * NT + threadIdx.x;
    rec[i] = op.load(idx < live ? idx : live - 1);
  }
#pragma unroll
  for (int i = 0; i < ITEMS; i++) {
    long long idx = base + (long long)i * NT + threadIdx.x;
    const int cnt = op.count(rec[i]);
    tile[i * NT + threadIdx.x] = idx < live ? cnt : 0;
  }
#pragma unroll
  for (int i = 0; i < ITEMS; i++) {
    long long idx = base + (long long)i * NT + threadIdx.x;
    if (idx < live) op.gather(idx, rec[i]);
  }
  __syncthreads();
  int loc[ITEMS];
  int s = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; i++) { loc[i] = s; s += tile[threadIdx.x * ITEMS + i]; }
  unsigned utot;
  int off = (int)block_exclusive_scan_nt<NT>((unsigned)s, lds, utot);
  const int tot = (int)utot;
  if (threadIdx.x == 0) {
    long long ex = 0;
    if (gridDim.x > 1) {   // one tile needs no look-back
      if (t == 0) {
        __hip_atomic_store(&desc[0], scan_word(epoch, SCAN_ST_PREFIX, tot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(&desc[t], scan_word(epoch, SCAN_ST_AGG, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long long j = t - 1;
        long long spins = 0;
        while (j >= 0) {
          unsigned long long w = __hip_atomic_load(&desc[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((unsigned)(w >> 42) != epoch) {
            if (++spins > (1LL << 20)) { atomicAdd(err, 1); break; }
            continue;
          }
          ex += (long long)(w & ((1ull << 40) - 1));
          if (((w >> 40) & 3ull) == SCAN_ST_PREFIX) break;
          j--;
        }
        __hip_atomic_store(&desc[t], scan_word(epoch, SCAN_ST_PREFIX, ex + tot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    excl_s = ex;
    if (total && base + TILE >= live) *total = (int)(ex + tot);
  }
  __syncthreads();
  off += (int)excl_s;
#pragma unroll
  for (int i = 0; i < ITEMS; i++) tile[threadIdx.x * ITEMS + i] = off + loc[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; i++) {
    long long idx = base + (long long)i * NT + threadIdx.x;
    if (idx < live) op.emit(idx, rec[i], tile[i * NT + threadIdx.x]);
  }
}
// the plain scan of an int array, and optionally the compaction of 0/1
// flags: the index of every set entry at its position in list
struct ScanArray {
  const int* in;
  int* out;
  int* list;
  struct Rec { int x; };
  __device__ Rec load(long long i) const { return {in[i]}; }
  __device__ int count(const Rec& r) const { return r.x; }
  __device__ void gather(long long, Rec&) const {}
  __device__ void emit(long long i, const Rec& r, int pos) const {
    out[i] = pos;
    if (list && r.x) list[pos] = (int)i;
  }
};

// ------------------------------------------------------------ step 1
// cusimp_free.cu:869-878 (stale pts_map remap) and :880-889 (table)
__global__ void k_remap_undo(const int* __restrict__ tmp, const int* __restrict__ pts_map, long long map_n,
                             int* __restrict__ undo_list, int first, int count, int* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int x = tmp[i];
  int r = 0;
  if (x >= 0 && (long long)x < map_n) r = pts_map[x];
  else atomicAdd(err, 1);
  undo_list[first + i] = r;
}
__global__ void k_set_table(const int* __restrict__ undo_in, int n, uint8_t* table, long long table_n, int* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x = undo_in[i];
  if (x >= 0 && (long long)x < table_n) table[x] = 1;
  else atomicAdd(err, 1);
}

// input validation: every face is (-1,-1,-1) or three indices in [0, V)
__global__ void k_validate(const int* __restrict__ T, int F, int V, int* err) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  I3 t = ldt(T, f);
  if (dead3(t)) return;
  if (t.i < 0 || t.i >= V || t.j < 0 || t.j >= V || t.k < 0 || t.k >= V) atomicAdd(err, 1);
}

__global__ void k_init_points(const float* __restrict__ pin, float* __restrict__ P, float* __restrict__ P0,
                              int* __restrict__ occ, int V) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > V) return;
  if (v == V) { occ[V] = 0; return; }
  float x = pin[3 * v], y = pin[3 * v + 1], z = pin[3 * v + 2];
  P[3 * v] = x; P[3 * v + 1] = y; P[3 * v + 2] = z;
  P0[3 * v] = x; P0[3 * v + 1] = y; P0[3 * v + 2] = z;
  occ[v] = 1;
}
__global__ void k_init_tris(const int* __restrict__ tin, int* __restrict__ T, int* __restrict__ T0, int F) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3LL * F) return;
  int x = tin[i];
  T[i] = x; T0[i] = x;
}

// ------------------------------------------------------------- shuffle
// thrust/rocThrust shuffle: Feistel bijection over the next power of two,
// keep keys < m in index order (thrust/system/detail/generic/shuffle.inl).
struct Feistel {
  uint32_t lbits, rbits, lmask, rmask;
  uint32_t key[24];
};
__device__ __forceinline__ uint64_t feistel(const Feistel& fb, uint64_t val) {
  uint32_t s0 = (uint32_t)(val >> fb.rbits), s1 = (uint32_t)(val & fb.rmask);
#pragma unroll 4
  for (int i = 0; i < 24; i++) {
    uint64_t prod = 0xD2B74407B1CE6E93ull * (uint64_t)s0;
    uint32_t hi = (uint32_t)(prod >> 32), lo = (uint32_t)prod;
    lo = (lo << (fb.rbits - fb.lbits)) | (s1 >> fb.lbits);
    s0 = ((hi ^ fb.key[i]) ^ s1) & fb.lmask;
    s1 = lo & fb.rmask;
  }
  return ((uint64_t)s0 << fb.rbits) | (uint64_t)s1;
}
__global__ void k_shuffle_flags(Feistel fb, long long n, long long m, int* flags) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = feistel(fb, (uint64_t)i) < (uint64_t)m ? 1 : 0;
}
__global__ void k_shuffle_scatter(Feistel fb, long long n, long long m, const int* __restrict__ pos,
                                  const int* __restrict__ tin, int* __restrict__ T, int* __restrict__ T0) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = feistel(fb, (uint64_t)i);
  if (k >= (uint64_t)m) return;
  int o = pos[i];
  for (int c = 0; c < 3; c++) {
    int x = tin[3 * k + c];
    T[3LL * o + c] = x; T0[3LL * o + c] = x;
  }
}

// ---------------------------------------------------------------- fans
// cusimp_free.cu:276-304 with canonical ascending-face order (D0)
__global__ void k_fan_count(const int* __restrict__ T, int F, int* cnt) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  I3 t = ldt(T, f);
  if (dead3(t)) return;
  atomicAdd(&cnt[t.i], 1); atomicAdd(&cnt[t.j], 1); atomicAdd(&cnt[t.k], 1);
}
__global__ void k_fan_fill(const int* __restrict__ T, int F, const int* __restrict__ fan_off, int* cursor,
                           int* __restrict__ fan) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  I3 t = ldt(T, f);
  if (dead3(t)) return;
  fan[fan_off[t.i] + atomicAdd(&cursor[t.i], 1)] = f;
  fan[fan_off[t.j] + atomicAdd(&cursor[t.j], 1)] = f;
  fan[fan_off[t.k] + atomicAdd(&cursor[t.k], 1)] = f;
}
__global__ void k_fan_sort(const int* __restrict__ fan_off, int* __restrict__ fan, int V) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  int a = fan_off[v], b = fan_off[v + 1];
  for (int i = a + 1; i < b; i++) {
    int x = fan[i];
    int j = i - 1;
    while (j >= a && fan[j] > x) { fan[j + 1] = fan[j]; j--; }
    fan[j + 1] = x;
  }
}

// --------------------------------------------------------------- edges
// cusimp_free.cu:306-335
__global__ void k_edge_count(const int* __restrict__ T, int F, int* __restrict__ cnt) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f > F) return;
  if (f == F) { cnt[F] = 0; return; }
  I3 t = ldt(T, f);
  cnt[f] = dead3(t) ? 0 : (int)(t.i > t.j) + (int)(t.j > t.k) + (int)(t.k > t.i);
}
__global__ void k_edge_emit(const int* __restrict__ T, int F, const int* __restrict__ eoff, int* __restrict__ edges) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  I3 t = ldt(T, f);
  if (dead3(t)) return;
  int o = eoff[f];
  if (t.i > t.j) { edges[2 * o] = t.j; edges[2 * o + 1] = t.i; o++; }
  if (t.j > t.k) { edges[2 * o] = t.k; edges[2 * o + 1] = t.j; o++; }
  if (t.k > t.i) { edges[2 * o] = t.i; edges[2 * o + 1] = t.k; o++; }
}

// ------------------------------------------------------------ quadrics
// cusimp_free.cu:337-366.  Q is symmetric (p_r*p_c == p_c*p_r exactly), so
// the 10 upper-triangle sums are stored: 00 01 02 03 11 12 13 22 23 33.
__global__ void __launch_bounds__(BLOCK) k_quadric(const float* __restrict__ P, const int* __restrict__ T,
                                                   const int* __restrict__ fan_off, const int* __restrict__ fan,
                                                   int V, float* __restrict__ Q) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  float q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int a = fan_off[v], b = fan_off[v + 1];
  for (int s = a; s < b; s++) {
    I3 t = ldt(T, fan[s]);
    V3 v0 = ldp(P, t.i), v1 = ldp(P, t.j), v2 = ldp(P, t.k);
    V3 n = vcross(vsub(v1, v0), vsub(v2, v0));
    n = vdivs(n, vnorm(n));
    float off = -vdot(n, v0);
    q[0] += n.x * n.x; q[1] += n.x * n.y; q[2] += n.x * n.z; q[3] += n.x * off;
    q[4] += n.y * n.y; q[5] += n.y * n.z; q[6] += n.y * off;
    q[7] += n.z * n.z; q[8] += n.z * off;
    q[9] += off * off;
  }
#pragma unroll
  for (int i = 0; i < 10; i++) Q[(long long)10 * v + i] = q[i];
}

// ----------------------------------------------------------- edge cost
__device__ __forceinline__ int succ_of(I3 t, int x) {
  if (t.i == x) return t.j;
  if (t.j == x) return t.k;
  if (t.k == x) return t.i;
  return -2;
}
__device__ __forceinline__ float tri_area(V3 p0, V3 p1, V3 p2) {   // cusimp_free.cu:368-375
  float a = vnorm(vsub(p0, p1)), b = vnorm(vsub(p1, p2)), c = vnorm(vsub(p2, p0));
  float s = (a + b + c) / 2;
  return sqrtf(s * (s - a) * (s - b) * (s - c));
}
// one skinny/flip pass, cusimp_free.cu:497-575; false when a normal flips.
// sp: the unguarded variant's term (cusimp.cu:423,463), where sqrt(3) takes
// an integer and so is the double overload: the whole quotient is evaluated
// in double and rounded to float once (D10).
__device__ bool skinny_pass(const float* __restrict__ P, const int* __restrict__ T, const int* __restrict__ fan_off,
                            const int* __restrict__ fan, int moved, int other, V3 m, float& Qa, int& nt, int sp = 0) {
  int a = fan_off[moved], b = fan_off[moved + 1];
  for (int s = a; s < b; s++) {
    I3 t = ldt(T, fan[s]);
    if (t.i == other || t.j == other || t.k == other) continue;
    V3 o0 = ldp(P, t.i), o1 = ldp(P, t.j), o2 = ldp(P, t.k);
    V3 n0 = t.i == moved ? m : o0, n1 = t.j == moved ? m : o1, n2 = t.k == moved ? m : o2;
    V3 on = vcross(vsub(o1, o0), vsub(o2, o0));
    on = vdivs(on, vnorm(on));
    V3 nn = vcross(vsub(n1, n0), vsub(n2, n0));
    nn = vdivs(nn, vnorm(nn));
    if (vdot(on, nn) < 0) return false;
    float e01 = vnorm(vsub(n0, n1)), e12 = vnorm(vsub(n1, n2)), e20 = vnorm(vsub(n2, n0));
    float term;
    if (sp) {
      const double k43d = 4.0 * 1.7320508075688772;   // (double)4.0f * sqrt(3.0), exact product
      term = (float)(k43d * (double)tri_area(n0, n1, n2) / (double)(e01 * e01) + (double)(e12 * e12) +
                     (double)(e20 * e20) + (double)0.0000001f);
    } else {
      const float k43 = 4.0f * 1.73205077648162841796875f;   // 4.0f * sqrtf(3.0f)
      term = k43 * tri_area(n0, n1, n2) / (e01 * e01) + e12 * e12 + e20 * e20 + 0.0000001f;
    }
    Qa += 1.0f - clamp_ref(term, 0.0f, 1.0f);
    nt++;
  }
  return true;
}
__device__ __forceinline__ float fan_edge_len(const float* __restrict__ P, const int* __restrict__ T,
                                              const int* __restrict__ fan_off, const int* __restrict__ fan, int w,
                                              float el, int& ne) {
  int a = fan_off[w], b = fan_off[w + 1];
  for (int s = a; s < b; s++) {
    I3 t = ldt(T, fan[s]);
    if (t.i > t.j) { el += vnorm(vsub(ldp(P, t.i), ldp(P, t.j))); ne++; }
    if (t.j > t.k) { el += vnorm(vsub(ldp(P, t.j), ldp(P, t.k))); ne++; }
    if (t.k > t.i) { el += vnorm(vsub(ldp(P, t.k), ldp(P, t.i))); ne++; }
  }
  return el;
}

// cusimp_free.cu:382-585; sp: the unguarded variant (cusimp.cu:283-482):
// no invalid-vertex table or stuck mode, the fan length term counted twice
// (:378), no x5 weight on the skinny term, the skinny term in double
__global__ void __launch_bounds__(BLOCK) k_edge_cost(const float* __restrict__ P, const int* __restrict__ T,
                                                     const int* __restrict__ fan_off, const int* __restrict__ fan,
                                                     const int* __restrict__ edges, int E,
                                                     const float* __restrict__ Q, const uint8_t* __restrict__ table,
                                                     uint32_t* __restrict__ cost, float scale, float tau, int is_stuck,
                                                     int sp = 0) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int u = edges[2 * e], v = edges[2 * e + 1];
  if (!sp) {
    if (table[u] && table[v]) { cost[e] = 0xFFFFFFFFu; return; }
    if (is_stuck) return;
  }
  V3 v0 = ldp(P, u), v1 = ldp(P, v);
  int ua = fan_off[u], ub = fan_off[u + 1], va = fan_off[v], vb = fan_off[v + 1];
  int dup = 0;
  for (int a = ua; a < ub; a++) {
    int iu = succ_of(ldt(T, fan[a]), u);
    for (int b = va; b < vb; b++) {
      int iv = succ_of(ldt(T, fan[b]), v);
      if (iu == iv) dup++;
      if (dup > 2) { cost[e] = 0xFFFFFFFFu; return; }
    }
  }
  if (dup != 2) return;
  int ne = 0;
  float el = fan_edge_len(P, T, fan_off, fan, u, 0.0f, ne);
  el = fan_edge_len(P, T, fan_off, fan, v, el, ne);
  el = el / (float)ne / scale * tau;
  if (sp) el = el + (el + vnorm(vsub(v0, v1)) / scale * tau);   // edge_len += edge_len + ...
  else el = el + vnorm(vsub(v0, v1)) / scale * tau;
  V3 m = vdivs(vadd(v0, v1), 2.0f);
  float Qa = 0.0f;
  int nt = 0;
  if (!skinny_pass(P, T, fan_off, fan, u, v, m, Qa, nt, sp)) { cost[e] = 0xFFFFFFFFu; return; }
  if (!skinny_pass(P, T, fan_off, fan, v, u, m, Qa, nt, sp)) { cost[e] = 0xFFFFFFFFu; return; }
  if (!sp) Qa = 5.0f * Qa;
  const float* qu = Q + (long long)10 * u;
  const float* qv = Q + (long long)10 * v;
  float s[10];
#pragma unroll
  for (int i = 0; i < 10; i++) s[i] = qu[i] + qv[i];
  // vTMv (cusimp_free.h:145-152) with M symmetric and w = (m, 1)
  float t0 = s[0] * m.x + s[1] * m.y + s[2] * m.z + s[3] * 1.0f;
  float t1 = s[1] * m.x + s[4] * m.y + s[5] * m.z + s[6] * 1.0f;
  float t2 = s[2] * m.x + s[5] * m.y + s[7] * m.z + s[8] * 1.0f;
  float t3 = s[3] * m.x + s[6] * m.y + s[8] * m.z + s[9] * 1.0f;
  float qe = t0 * m.x + t1 * m.y + t2 * m.z + t3 * 1.0f;
  float c = qe / (scale * scale);
  cost[e] = u32sat(clamp_ref(c + el + Qa / (float)nt * tau, 0.0f, COST_RANGE) / COST_RANGE * U32F);
}

// ---------------------------------------------------------------- claim
// cusimp_free.cu:587-610
__global__ void __launch_bounds__(BLOCK) k_claim(const int* __restrict__ edges, int E, const uint32_t* __restrict__ cost,
                                                 const int* __restrict__ fan_off, const int* __restrict__ fan,
                                                 u64* tri_min) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  u64 key = ((u64)cost[e] << 32) | (u64)(uint32_t)e;
  for (int w = 0; w < 2; w++) {
    int x = edges[2 * e + w];
    int a = fan_off[x], b = fan_off[x + 1];
    for (int s = a; s < b; s++) atomicMin(&tri_min[fan[s]], key);
  }
}

// ------------------------------------------------------------- collapse
__device__ __forceinline__ bool is_skinny(V3 p, V3 q, V3 r) {   // cusimp_free.cu:617-637
  V3 pq = vsub(q, p), pr = vsub(r, p), qr = vsub(r, q);
  float lpq = sqrtf(vdot(pq, pq)), lpr = sqrtf(vdot(pr, pr)), lqr = sqrtf(vdot(qr, qr));
  float ml = fmaxf(lpq, fmaxf(lpr, lqr));
  V3 c = vcross(pq, pr);
  float thr = (float)(1e-6 * (double)ml * (double)ml);
  float area = 0.5f * sqrtf(c.x * c.x + c.y * c.y + c.z * c.z);
  return area < thr;
}
__device__ __forceinline__ void surgery(int* T, const int* __restrict__ fan_off, const int* __restrict__ fan, int u,
                                        int v) {
  for (int s = fan_off[u]; s < fan_off[u + 1]; s++) {
    int f = fan[s];
    I3 t = ldt(T, f);
    if (t.i == v || t.j == v || t.k == v) { T[3 * f] = -1; T[3 * f + 1] = -1; T[3 * f + 2] = -1; }
  }
  for (int s = fan_off[v]; s < fan_off[v + 1]; s++) {
    int f = fan[s];
    I3 t = ldt(T, f);
    if (t.i == u || t.j == u || t.k == u) { T[3 * f] = -1; T[3 * f + 1] = -1; T[3 * f + 2] = -1; }
    else if (t.i == v) T[3 * f] = u;
    else if (t.j == v) T[3 * f + 1] = u;
    else if (t.k == v) T[3 * f + 2] = u;
  }
}
__device__ __forceinline__ void restore_edge(int u, int v, float* P, const float* __restrict__ P0, int* T,
                                             const int* __restrict__ T0, int* occ, const int* __restrict__ fan_off,
                                             const int* __restrict__ fan) {
  occ[v] = 1;
  stp(P, u, ldp(P0, u));
  stp(P, v, ldp(P0, v));
  for (int w = 0; w < 2; w++) {
    int x = w == 0 ? u : v;
    for (int s = fan_off[x]; s < fan_off[x + 1]; s++) {
      int f = fan[s];
      T[3 * f] = T0[3 * f]; T[3 * f + 1] = T0[3 * f + 1]; T[3 * f + 2] = T0[3 * f + 2];
    }
  }
}
// cusimp_free.cu:639-774.  Winners own disjoint fans and one-rings, so the
// collapse is applied by the winner's own thread; win[e] flags the list.
__global__ void __launch_bounds__(BLOCK) k_collapse(const int* __restrict__ edges, int E, const uint32_t* __restrict__ cost,
                                                    uint32_t collapse_t, const int* __restrict__ fan_off,
                                                    const int* __restrict__ fan, const u64* __restrict__ tri_min,
                                                    float* P, const float* __restrict__ P0, int* T,
                                                    const int* __restrict__ T0, int* occ, int* __restrict__ win,
                                                    int sp = 0) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  uint32_t c = cost[e];
  int u = edges[2 * e], v = edges[2 * e + 1];
  bool ok = c <= collapse_t;
  if (ok) {
    u64 key = ((u64)c << 32) | (u64)(uint32_t)e;
    for (int w = 0; w < 2 && ok; w++) {
      int x = w == 0 ? u : v;
      for (int s = fan_off[x]; s < fan_off[x + 1]; s++)
        if (tri_min[fan[s]] != key) { ok = false; break; }
    }
  }
  win[e] = ok ? 1 : 0;
  if (!ok) return;
  V3 m = vdivs(vadd(ldp(P, u), ldp(P, v)), 2.0f);
  stp(P, u, m);
  stp(P, v, V3{0.0f, 0.0f, 0.0f});
  occ[v] = 0;
  surgery(T, fan_off, fan, u, v);
  if (sp) return;   // the unguarded variant keeps every winner (cusimp.cu:511-584)
  bool sk = false;
  for (int w = 0; w < 2 && !sk; w++) {
    int x = w == 0 ? u : v;
    for (int s = fan_off[x]; s < fan_off[x + 1]; s++) {
      I3 t = ldt(T, fan[s]);
      if (t.i == -1) continue;
      if (is_skinny(ldp(P, t.i), ldp(P, t.j), ldp(P, t.k))) { sk = true; break; }
    }
  }
  if (sk) restore_edge(u, v, P, P0, T, T0, occ, fan_off, fan);
}

// compaction of a 0/1 flag array into ascending indices (pos = excl. scan)
__global__ void k_compact_idx(const int* __restrict__ flags, const int* __restrict__ pos, long long n,
                              int* __restrict__ out, int out_base) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i]) out[out_base + pos[i]] = (int)i;
}

// cusimp_free.cu:218-232
__global__ void k_remove_invalid_faces(int* T, int F) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  I3 t = ldt(T, f);
  if (dead3(t)) return;
  if (t.i == t.j || t.j == t.k || t.k == t.i) { T[3 * f] = -1; T[3 * f + 1] = -1; T[3 * f + 2] = -1; }
}

// cusimp_free.cu:234-274: detect (parallel) ...
__global__ void k_line_detect(const float* __restrict__ P, const int* __restrict__ edges, int E, int* count) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  V3 p = ldp(P, edges[2 * e]), q = ldp(P, edges[2 * e + 1]);
  if (p.x == q.x && p.y == q.y && p.z == q.z) atomicAdd(count, 1);
}
// ... and apply sequentially in ascending edge order (D5); one thread.
__global__ void k_line_apply_seq(float* P, int* T, int* occ, const int* __restrict__ edges, int E,
                                 const int* __restrict__ fan_off, const int* __restrict__ fan, int* collapsed,
                                 int* n_collapsed) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int n = *n_collapsed;
  for (int e = 0; e < E; e++) {
    int u = edges[2 * e], v = edges[2 * e + 1];
    V3 p = ldp(P, u), q = ldp(P, v);
    if (!(p.x == q.x && p.y == q.y && p.z == q.z)) continue;
    collapsed[n++] = e;
    stp(P, v, V3{0.0f, 0.0f, 0.0f});
    occ[v] = 0;
    surgery(T, fan_off, fan, u, v);
  }
  *n_collapsed = n;
}

// ------------------------------------------------------- undo support
// faces in the fans of the first `cut` collapsed edges -> qflag
__global__ void k_qflag_mark(const int* __restrict__ collapsed, const int* __restrict__ cut_ptr,
                             const int* __restrict__ edges, const int* __restrict__ fan_off,
                             const int* __restrict__ fan, uint8_t* qflag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *cut_ptr) return;
  int e = collapsed[i];
  for (int w = 0; w < 2; w++) {
    int x = edges[2 * e + w];
    for (int s = fan_off[x]; s < fan_off[x + 1]; s++) qflag[fan[s]] = 1;
  }
}
__global__ void k_query_flags(const uint8_t* __restrict__ qflag, const int* __restrict__ T, int F,
                              int* __restrict__ qf, int* __restrict__ af) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  bool alive = T[3 * f] != -1 || T[3 * f + 1] != -1 || T[3 * f + 2] != -1;
  af[f] = alive ? 1 : 0;
  qf[f] = (alive && qflag[f]) ? 1 : 0;
}
// cusimp_free.cu:776-827 with a face bitmap instead of a list scan
__global__ void k_mark_undo(const int* __restrict__ collapsed, const int* __restrict__ cut_ptr,
                            const int* __restrict__ edges, const int* __restrict__ fan_off,
                            const int* __restrict__ fan, const uint8_t* __restrict__ S, int* __restrict__ mflag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cut = *cut_ptr;
  if (i >= cut) return;
  int e = collapsed[i];
  int hit = 0;
  for (int w = 0; w < 2 && !hit; w++) {
    int x = edges[2 * e + w];
    for (int s = fan_off[x]; s < fan_off[x + 1]; s++) if (S[fan[s]]) { hit = 1; break; }
  }
  mflag[i] = hit;
}
// marked list (positions in the collapsed list, ascending) -> edge ids and
// endpoint pairs (tmp_vertices_undo_list, cusimp_free.cu:795-798)
__global__ void k_marked_pairs(const int* __restrict__ mflag, const int* __restrict__ mpos,
                               const int* __restrict__ cut_ptr, const int* __restrict__ collapsed,
                               const int* __restrict__ edges, int* __restrict__ marked, int* __restrict__ tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *cut_ptr) return;
  if (!mflag[i]) return;
  int j = mpos[i];
  int e = collapsed[i];
  marked[j] = e;
  tmp[2 * j] = edges[2 * e];
  tmp[2 * j + 1] = edges[2 * e + 1];
}
// cusimp_free.cu:829-867
__global__ void k_undo(const int* __restrict__ marked, const int* __restrict__ n_ptr, const int* __restrict__ edges,
                       float* P, const float* __restrict__ P0, int* T, const int* __restrict__ T0, int* occ,
                       const int* __restrict__ fan_off, const int* __restrict__ fan) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *n_ptr) return;
  int e = marked[i];
  restore_edge(edges[2 * e], edges[2 * e + 1], P, P0, T, T0, occ, fan_off, fan);
}

// ----------------------------------------------------------- compaction
// PaMO.run between rounds (simp_cuda/pamo/__init__.py:136-140)
__global__ void k_compact_verts(const float* __restrict__ P, const int* __restrict__ occ,
                                const int* __restrict__ pts_map, int V, float* __restrict__ out) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  if (!occ[v]) return;
  int o = pts_map[v];
  out[3 * o] = P[3 * v]; out[3 * o + 1] = P[3 * v + 1]; out[3 * o + 2] = P[3 * v + 2];
}
__global__ void k_face_keep(const int* __restrict__ T, int F, int* __restrict__ keep) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  keep[f] = T[3 * f] >= 0 ? 1 : 0;
}
__global__ void k_compact_faces(const int* __restrict__ T, const int* __restrict__ keep, const int* __restrict__ pos,
                                const int* __restrict__ pts_map, int F, int* __restrict__ out) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  if (!keep[f]) return;
  int o = pos[f];
  out[3 * o] = pts_map[T[3 * f]]; out[3 * o + 1] = pts_map[T[3 * f + 1]]; out[3 * o + 2] = pts_map[T[3 * f + 2]];
}

}  // namespace pamo
