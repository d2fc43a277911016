// Kernels of the batched engine: many meshes packed CSR-style in one set of
// buffers and simplified together, one launch per step for all of them.
//
// Packing: vertices and faces of mesh k occupy [voff[k], voff[k+1]) and
// [foff[k], foff[k+1]); face indices are global.  vmesh/fmesh give the mesh
// of a vertex/face.  Persistent per-mesh state (the previous round's vertex
// map, the invalid-vertex table and the edge costs, which the reference keeps
// between calls, simp_cuda/src/cusimp_free.h:184-280) lives in per-mesh
// regions at fixed offsets (pmo[k] for map/table, eco[k] for edge costs)
// addressed by LOCAL indices, exactly as one reference context per mesh.
//
// Edge indices are global and ascending within a mesh, so the claim keys
// (cost << 32 | edge) order the edges of one mesh exactly as the reference's
// local indices do (keys of different meshes are never compared: a face
// belongs to one mesh).
#pragma once
#include "pamo_bvh.h"
#include "pamo_kernels.h"
#include "placement.h"
#include "flip_filter.h"

namespace pamo {
namespace eng {

// per-mesh parameters of a round (device copy of a host struct array)
struct MeshParam {
  float scale, tau;
  uint32_t ct;            // quantized collapse threshold
  int stuck;
  long long pmo, eco;     // persistent region offsets (map/table, edge costs)
  long long pmcap;        // map/table region entries (alloc_pts + 1)
  float box[6];           // vertex bounds at admission (lo xyz, hi xyz): every
                          // later position is a midpoint of earlier ones or a
                          // restored original, so the box bounds the mesh
  long long target, iters;   // termination (simp_cuda/pamo/__init__.py:142-158)
  int tolerance, pad;
};

// A mesh's run state between rounds, kept on the device by the round's
// status kernel: its sizes after the last round and the termination
// counters of PaMO.run (simp_cuda/pamo/__init__.py:142-158).
struct MeshState {
  int F, V, n_undo;   // faces, vertices and undo-list length after the last round
  int count, stuck;   // rounds without a change of F; the is_stuck flag
  int rounds;         // rounds run
  int exit;           // 0 running, 1 target, 2 <= 10 faces, 3 tolerance, 4 iteration cap
  int done;
};
// Words of the engine's live-size buffer: the next round's totals (the
// round reads them on the device, so a round can be enqueued before the host
// knows them), the attention flag and the count of meshes the last round
// finished.  A non-zero attention flag skips every launch of a round queued
// behind it (gated()): the last round marked collapses for undo, or
// finished a mesh, and the host acts before anything else runs.
// The epoch word is the base of the scans' device epochs in a replayed
// graph (k_scan_op); each round's status moves it on.
constexpr int LIVE_V = 0, LIVE_F = 1, LIVE_UIN = 2, LIVE_ATTN = 3, LIVE_DONE = 4, LIVE_EPOCH = 5, LIVE_WORDS = 8;

// Undo passes per round: the reference repeats the check at most 5 times
// (cusimp_free.cu:1014-1040); the strict guard repeats until nothing is
// marked, at most STRICT_PASSES times (only a round whose input already
// intersects needs more than a few).
constexpr int REF_PASSES = 5;
constexpr int STRICT_PASSES = 64;
__host__ __device__ constexpr int undo_passes_cap(int strict) { return strict ? STRICT_PASSES : REF_PASSES; }
// undo-list entries a mesh may write per scanned collapsed entry in one
// round: 2 per marked entry for the first check and for every pass
__host__ __device__ constexpr long long undo_slots_per_cut(int strict) { return 2LL * (undo_passes_cap(strict) + 1); }

__device__ __forceinline__ int find_seg(const long long* __restrict__ off, int M, long long i) {
  int lo = 0, hi = M - 1;   // largest k with off[k] <= i
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------- step 1 (undo in)
// clear the table of meshes that bring an undo list and are not stuck
// (cusimp_free.cu:905-908); one block per mesh
// every incoming undo entry marks its mesh's table (cusimp_free.cu:910-916);
// bounds-checked (the reference is not)
__global__ void k_undo_table(const int* __restrict__ undo_in, const long long* __restrict__ uioff, int M,
                             const MeshParam* __restrict__ mp, uint8_t* __restrict__ table, int* err,
                             const int* __restrict__ gate) {
  if (gated(gate)) return;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= uioff[M]) return;
  int k = find_seg(uioff, M, i);
  int x = undo_in[i];
  if (x >= 0 && (long long)x < mp[k].pmcap) table[mp[k].pmo + x] = 1;
  else atomicAdd(err, 1);
}
// stuck meshes copy their incoming list through the stale map into the head
// of their undo output (cusimp_free.cu:899-904)
__global__ void k_undo_remap_in(const int* __restrict__ undo_in, const long long* __restrict__ uioff, int M,
                                const MeshParam* __restrict__ mp, const int* __restrict__ pm,
                                const long long* __restrict__ uobase, int* __restrict__ undo_out,
                                const int* __restrict__ gate) {
  if (gated(gate)) return;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= uioff[M]) return;
  int k = find_seg(uioff, M, i);
  if (!mp[k].stuck) return;
  int x = undo_in[i];
  bool ok = x >= 0 && (long long)x < mp[k].pmcap;
  undo_out[uobase[k] + (i - uioff[k])] = ok ? pm[mp[k].pmo + x] : 0;
}
__global__ void k_init_mesh_state(int M, const int* __restrict__ nust, int* __restrict__ first, int* __restrict__ nvu,
                                  int* __restrict__ lcnt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  first[k] = nust[k];
  nvu[k] = nust[k];
  lcnt[k] = 0;
}

// Round start in one launch: counters (with the live V and F), round
// statistics, per-mesh undo state, the three per-vertex counters of the fan
// and edge builds, the occupancy flags (1 per vertex, 0 at V) and the
// per-vertex claim minima.
// It also clears the invalid-vertex table of every mesh that receives an
// undo list and is not stuck (cusimp_free.cu:905-907), one entry per vertex:
// the table is read at local vertex indices, which stay below the mesh's
// current vertex count for the rest of the run (it never grows), so entries
// past it are never read again.
__global__ void k_round_init(const int* __restrict__ live, int M, int* __restrict__ counters, int n_counters,
                             int c_v, int c_f, unsigned* __restrict__ cstats, int n_cstats,
                             const int* __restrict__ nust, int* __restrict__ first, int* __restrict__ nvu,
                             int* __restrict__ lcnt, int* __restrict__ cnt, int* __restrict__ ucnt,
                             int* __restrict__ occ, u64* __restrict__ vmin, const int* __restrict__ vmesh,
                             const MeshParam* __restrict__ mp, const long long* __restrict__ uioff,
                             const long long* __restrict__ voff, uint8_t* __restrict__ table,
                             int* __restrict__ npass, const int* __restrict__ gate) {
  if (gated(gate)) return;
  const long long V = live[LIVE_V], F = live[LIVE_F];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < V) {
    const int k = vmesh[i];
    if (!mp[k].stuck && uioff[k + 1] != uioff[k]) table[mp[k].pmo + (i - voff[k])] = 0;
  }
  if (i < n_counters) counters[i] = i == c_v ? (int)V : (i == c_f ? (int)F : 0);
  if (i < n_cstats) cstats[i] = 0;
  if (i < M) {
    first[i] = nust[i];
    nvu[i] = nust[i];
    lcnt[i] = 0;
    npass[i] = 0;
  }
  if (i <= V) {
    cnt[i] = 0;
    ucnt[i] = 0;
    occ[i] = i < V ? 1 : 0;
    if (i < V) vmin[i] = ~0ull;
  }
}
// fan counts, plus the round-start face snapshot (the undo restore source).
// The count's atomic returns the face's slot in each corner's fan (in an
// arbitrary order: k_fan_slots sorts every fan), kept in frank so the fill
// is plain stores.
__global__ void k_fan_count_snap(const int* __restrict__ T, const int* __restrict__ F_ptr, int* cnt,
                                 int* __restrict__ T0, int* __restrict__ frank, const int* __restrict__ gate) {
  if (gated(gate)) return;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= *F_ptr) return;
  I3 t = ldt(T, f);
  T0[3LL * f] = t.i;
  T0[3LL * f + 1] = t.j;
  T0[3LL * f + 2] = t.k;
  if (dead3(t)) return;
  const int ri = atomicAdd(&cnt[t.i], 1), rj = atomicAdd(&cnt[t.j], 1), rk = atomicAdd(&cnt[t.k], 1);
  frank[3LL * f] = ri;
  frank[3LL * f + 1] = rj;
  frank[3LL * f + 2] = rk;
}
// the fans from the counts' slots
__global__ void k_fan_place(const int* __restrict__ T, const int* __restrict__ F_ptr,
                            const int* __restrict__ fan_off, const int* __restrict__ frank, int* __restrict__ fan,
                            const int* __restrict__ gate) {
  if (gated(gate)) return;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= *F_ptr) return;
  I3 t = ldt(T, f);
  if (dead3(t)) return;
  const I3 r = ldt(frank, f);
  fan[fan_off[t.i] + r.i] = f;
  fan[fan_off[t.j] + r.j] = f;
  fan[fan_off[t.k] + r.k] = f;
}
// Per-check resets in one launch: query and hit flags of every face, the
// per-mesh leaf ranges, the root's parent, the mark counter and the pair
// buffer counters, and the overflow flags of the exhaustive search.
__global__ void k_check_init(long long F, int M, uint8_t* __restrict__ qflag, uint8_t* __restrict__ S,
                             int* __restrict__ seg_a, int* __restrict__ seg_b, int* __restrict__ kparent,
                             int* __restrict__ nmark, int* __restrict__ novf, unsigned long long* __restrict__ npairs,
                             int* __restrict__ oflag) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F) {
    qflag[i] = 0;
    S[i] = 0;
    oflag[i] = 0;
  }
  if (i < M) {
    seg_a[i] = -1;
    seg_b[i] = -1;
  }
  if (i == 0) {
    kparent[0] = -1;
    *nmark = 0;
    *novf = 0;
    *npairs = 0;
  }
}

// admission: faces with local indices -> global, with or without the shuffle
__global__ void k_copy_faces_off(const int* __restrict__ tin, int* __restrict__ T, int F, int voff) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3LL * F) return;
  T[i] = tin[i] + voff;
}
__global__ void k_shuffle_scatter_off(Feistel fb, long long n, long long m, const int* __restrict__ pos,
                                      const int* __restrict__ tin, int* __restrict__ T, int voff) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = feistel(fb, (uint64_t)i);
  if (k >= (uint64_t)m) return;
  int o = pos[i];
  for (int c = 0; c < 3; c++) T[3LL * o + c] = tin[3 * k + c] + voff;
}

// ------------------------------------------------------ packed transfers
// up to COPY_SEGS word ranges copied in one launch (grid.y = range)
constexpr int COPY_SEGS = 8;
struct CopySegs {
  const int* src[COPY_SEGS];
  int* dst[COPY_SEGS];
  int nw[COPY_SEGS];
  int n;
};
__global__ void k_copy_segs(CopySegs cs) {
  const int q = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < cs.n && i < cs.nw[q]) cs.dst[q][i] = cs.src[q][i];
}

// ------------------------------------------------------ fan slots
// Per vertex: sorts its fan (ascending face index, D0) and writes, per fan
// slot, what the edge costs read about that face, so the cost kernel
// walks contiguous slot arrays instead of gathering face records:
//   fsucc  successor of the vertex in the face (link condition,
//          cusimp_free.cu:407-430)
//   sface  the face's vertex indices and its index
//   splane unit normal and offset (tri2plane, cusimp_free.cu:337-346; also
//          the old normal of the flip test :520-521)
//   slen   |Pi-Pj|, |Pj-Pk|, |Pk-Pi| (:451-483)
// and the vertex quadric from the same planes in fan order (:348-366);
// it also takes the round-start snapshot of the vertex position.
// A face's values are computed by each of its three vertices with the same
// expressions, so all copies are identical.
// the data of one slot: the face's corners and index, the successor of w,
// the plane and the edge lengths; returns the plane for the quadric
__device__ __forceinline__ float4 fan_slot_fill(const float* __restrict__ P, const int* __restrict__ T, int w, int f,
                                               int s, int* __restrict__ fsucc, int4* __restrict__ sface,
                                               float4* __restrict__ splane, float4* __restrict__ slen) {
  I3 t = ldt(T, f);
  fsucc[s] = succ_of(t, w);
  sface[s] = make_int4(t.i, t.j, t.k, f);
  V3 v0 = ldp(P, t.i), v1 = ldp(P, t.j), v2 = ldp(P, t.k);
  V3 n = vcross(vsub(v1, v0), vsub(v2, v0));
  n = vdivs(n, vnorm(n));
  float off = -vdot(n, v0);
  float4 pl = make_float4(n.x, n.y, n.z, off);
  splane[s] = pl;
  slen[s] = make_float4(vnorm(vsub(v0, v1)), vnorm(vsub(v1, v2)), vnorm(vsub(v2, v0)), 0.0f);
  return pl;
}
// Q = sum over the fan, in fan order, of p p^T (upper triangle)
__device__ __forceinline__ void quadric_add(float* q, float4 p) {
  q[0] += p.x * p.x; q[1] += p.x * p.y; q[2] += p.x * p.z; q[3] += p.x * p.w;
  q[4] += p.y * p.y; q[5] += p.y * p.z; q[6] += p.y * p.w;
  q[7] += p.z * p.z; q[8] += p.z * p.w;
  q[9] += p.w * p.w;
}

// A group of FS_G lanes takes one vertex.  The fan is sorted by rank in LDS
// (face ids in a fan are distinct, so the rank of x is the count of smaller
// entries), the lanes fill the slots in parallel, and one lane adds the
// planes into the quadric serially in fan order.  Fans longer than FS_MAX
// rank from global memory instead: each lane fills the slot at its entry's
// rank, and the sorted fan is read back from the slots' face indices.
constexpr int FS_G = 8;
constexpr int FS_MAX = 32;
constexpr int FS_GROUPS = BLOCK / FS_G;

__global__ void __launch_bounds__(BLOCK) k_fan_slots(const float* __restrict__ P, const int* __restrict__ T,
                                                     const int* __restrict__ fan_off, int* __restrict__ fan,
                                                     const int* __restrict__ V_ptr,
                                                     int* __restrict__ fsucc, int4* __restrict__ sface,
                                                     float4* __restrict__ splane, float4* __restrict__ slen,
                                                     float* __restrict__ Q, float* __restrict__ P0,
                                                     const int* __restrict__ gate) {
  if (gated(gate)) return;   // uniform over the workgroup
  const int V = *V_ptr;
  __shared__ int s_raw[FS_GROUPS][FS_MAX];
  __shared__ int s_sorted[FS_GROUPS][FS_MAX];
  __shared__ float4 s_plane[FS_GROUPS][FS_MAX];
  const int grp = threadIdx.x / FS_G, g = threadIdx.x & (FS_G - 1);
  const int v = blockIdx.x * FS_GROUPS + grp;
  const bool valid = v < V;
  int a = 0, deg = 0;
  if (valid) {
    a = fan_off[v];
    deg = fan_off[v + 1] - a;
  }
  const bool small = deg <= FS_MAX;
  if (valid && small)
    for (int s = g; s < deg; s += FS_G) s_raw[grp][s] = fan[a + s];
  __syncthreads();
  if (valid && small) {
    for (int s = g; s < deg; s += FS_G) {
      int x = s_raw[grp][s], r = 0;
      for (int j = 0; j < deg; j++) r += s_raw[grp][j] < x;
      s_sorted[grp][r] = x;
    }
  } else if (valid) {
    for (int s = g; s < deg; s += FS_G) {
      int x = fan[a + s], r = 0, j = 0;
      // eight loads in flight per step
      for (; j + 8 <= deg; j += 8) {
        int y[8];
#pragma unroll
        for (int q = 0; q < 8; q++) y[q] = fan[a + j + q];
#pragma unroll
        for (int q = 0; q < 8; q++) r += y[q] < x;
      }
      for (; j < deg; j++) r += fan[a + j] < x;
      fan_slot_fill(P, T, v, x, a + r, fsucc, sface, splane, slen);
    }
  }
  __syncthreads();   // also orders the long fans' slot writes before their reads below
  if (valid && small) {
    for (int s = g; s < deg; s += FS_G) {
      int f = s_sorted[grp][s];
      fan[a + s] = f;
      s_plane[grp][s] = fan_slot_fill(P, T, v, f, a + s, fsucc, sface, splane, slen);
    }
  } else if (valid) {
    for (int s = g; s < deg; s += FS_G) fan[a + s] = sface[a + s].w;
  }
  __syncthreads();
  if (!valid || g != 0) return;
  stp(P0, v, ldp(P, v));   // round-start position snapshot (the undo restore source)
  float q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (small) {
    for (int s = 0; s < deg; s++) quadric_add(q, s_plane[grp][s]);
  } else {
    int s = a;
    for (; s + 8 <= a + deg; s += 8) {
      float4 pl[8];
#pragma unroll
      for (int j = 0; j < 8; j++) pl[j] = splane[s + j];
#pragma unroll
      for (int j = 0; j < 8; j++) quadric_add(q, pl[j]);
    }
    for (; s < a + deg; s++) quadric_add(q, splane[s]);
  }
#pragma unroll
  for (int i = 0; i < 10; i++) Q[(long long)10 * v + i] = q[i];
}

// ------------------------------------------- scans with their producers
// Edge list (cusimp_free.cu:306-335): per face the count of its descending
// half-edges, scanned to the face's first edge; the emit writes the edges in
// face order and ranks each within its first endpoint (the rank places the
// edge in the vertex-ordered list, k_edge_vorder).
struct EdgeScanOp {
  const int* T;
  const int* F_ptr;
  int* eoff;
  int* edges;
  int* ucnt;
  int* erank;
  int* win;   // winner flags of the edges (and the entry past the last), cleared
  struct Rec { I3 t; int r0, r1, r2; };
  // f = F (the entry past the last face) reads face max(F - 1, 0), which
  // the buffer always holds, and takes it as dead
  __device__ Rec load(long long f) const {
    const long long F = *F_ptr;
    const I3 t = ldt(T, (int)max(min(f, F - 1), 0LL));
    Rec r;
    r.t = f < F ? t : I3{-1, -1, -1};
    return r;
  }
  __device__ int count(const Rec& r) const {
    const I3 t = r.t;
    return dead3(t) ? 0 : (int)(t.i > t.j) + (int)(t.j > t.k) + (int)(t.k > t.i);
  }
  // ranks within the first endpoints (dead faces have no descending edge)
  __device__ void gather(long long, Rec& r) const {
    const I3 t = r.t;
    if (t.i > t.j) r.r0 = atomicAdd(&ucnt[t.j], 1);
    if (t.j > t.k) r.r1 = atomicAdd(&ucnt[t.k], 1);
    if (t.k > t.i) r.r2 = atomicAdd(&ucnt[t.i], 1);
  }
  __device__ void emit(long long f, const Rec& r, int o) const {
    eoff[f] = o;
    if (f >= *F_ptr) { win[o] = 0; return; }
    const I3 t = r.t;
    if (t.i > t.j) { edges[2 * o] = t.j; edges[2 * o + 1] = t.i; erank[o] = r.r0; win[o++] = 0; }
    if (t.j > t.k) { edges[2 * o] = t.k; edges[2 * o + 1] = t.j; erank[o] = r.r1; win[o++] = 0; }
    if (t.k > t.i) { edges[2 * o] = t.i; edges[2 * o + 1] = t.k; erank[o] = r.r2; win[o++] = 0; }
  }
};
// End-of-round vertex compaction: the scan of occupancy is the new index g,
// and occupied vertices move to it with their mesh id.
struct VertCompactOp {
  const int* occ;
  int* g;
  const int* V_ptr;
  const float* P;
  const int* vmesh;
  float* outP;
  int* outm;
  struct Rec { int o; bool keep; V3 p; int m; };
  __device__ Rec load(long long v) const {
    Rec r;
    r.o = occ[v];
    r.keep = false;
    return r;
  }
  __device__ int count(const Rec& r) const { return r.o; }
  __device__ void gather(long long v, Rec& r) const {
    r.keep = v < *V_ptr && r.o;
    if (r.keep) {
      r.p = ldp(P, (int)v);
      r.m = vmesh[v];
    }
  }
  __device__ void emit(long long v, const Rec& r, int o) const {
    g[v] = o;
    if (!r.keep) return;
    stp(outP, o, r.p);
    outm[o] = r.m;
  }
};
// End-of-round face compaction: kept faces (not removed) move to the scan
// position with their corners renamed through g.
struct FaceCompactOp {
  const int* T;
  const int* F_ptr;
  int* kpos;
  const int* g;
  const int* fmesh;
  int* outT;
  int* outm;
  struct Rec { I3 t; int m; };
  // the same clamp as the edge list's
  __device__ Rec load(long long f) const {
    const long long Fd = *F_ptr;
    const I3 t = ldt(T, (int)max(min(f, Fd - 1), 0LL));
    Rec r;
    r.t = f < Fd ? t : I3{-1, -1, -1};
    return r;
  }
  __device__ int count(const Rec& r) const { return r.t.i >= 0 ? 1 : 0; }
  // the corners renamed through g
  __device__ void gather(long long f, Rec& r) const {
    if (r.t.i < 0) return;
    r.t = I3{g[r.t.i], g[r.t.j], g[r.t.k]};
    r.m = fmesh[f];
  }
  __device__ void emit(long long f, const Rec& r, int o) const {
    kpos[f] = o;
    if (r.t.i < 0) return;
    outT[3 * o] = r.t.i; outT[3 * o + 1] = r.t.j; outT[3 * o + 2] = r.t.k;
    outm[o] = r.m;
  }
};

// ------------------------------------------------- vertex-ordered edges
// Edges keep their index (the claim key's tie-break), but are processed
// grouped by their first endpoint u, so neighbouring lanes read the same
// fans.  The edge emit counts edges per u (the atomic's return value is
// the edge's rank within u; any rank order gives the same results, since
// every output depends only on the edge index).
// (threads 0..M also write each mesh's first edge, eoffm)
__global__ void k_edge_vorder(const int* __restrict__ edges, const int* __restrict__ E_ptr,
                              const int* __restrict__ uoff, const int* __restrict__ erank, int4* __restrict__ vedge,
                              int M, const long long* __restrict__ foff, const int* __restrict__ eoff_face,
                              long long* __restrict__ eoffm, const int* __restrict__ gate) {
  if (gated(gate)) return;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e <= M) eoffm[e] = eoff_face[foff[e]];
  if (e >= *E_ptr) return;
  int u = edges[2 * e], v = edges[2 * e + 1];
  vedge[uoff[u] + erank[e]] = make_int4(e, u, v, 0);
}

// The position an edge collapses to: the midpoint (the reference,
// cusimp_free.cu:491,687-688), or with placement 1 the optimal vertex of
// Q_u + Q_v (placement.h); the cost and the collapse call the same function.
__device__ __forceinline__ V3 edge_position(const float* __restrict__ Q, int u, int v, V3 pu, V3 pv, int placement) {
  if (!placement) return vdivs(vadd(pu, pv), 2.0f);
  const float* qu = Q + (long long)10 * u;
  const float* qv = Q + (long long)10 * v;
  float q[10];
#pragma unroll
  for (int j = 0; j < 10; j++) q[j] = qu[j] + qv[j];
  const place::P3 x = place::optimal(q, place::P3{pu.x, pu.y, pu.z}, place::P3{pv.x, pv.y, pv.z}, nullptr);
  return V3{x.x, x.y, x.z};
}

// ----------------------------------------------------------- edge cost
// cusimp_free.cu:382-585 with per-mesh parameters.  A group of EC_G lanes
// takes one edge (edges in vertex order); lane g loads slots g, g+EC_G, ...
// of each fan, so the loads of a fan issue together instead of as a chain
// of dependent loads per lane.  What depends on order stays serial: the
// edge-length sum and the skinny sum are accumulated term by term in fan
// order (fan u, then fan v) by every lane of the group from shuffled
// per-slot terms.  A term that the reference does not add is added as +0,
// which leaves the sum unchanged: both sums start at +0 and every term is
// +0 or more (a norm, or 1 - clamp(x, 0, 1)) or NaN.  The link count and the
// flip test are an integer sum and an OR, exact in any order; a flip
// anywhere makes the cost the maximum, as the reference's early return.
// EC_G is the default group size (4 lanes: 11.7 ms per 2.0M-face run,
// against 13.1 ms with 8, 14.6 ms with 2 and 20.3 ms with 16); the kernel is
// instantiated for 2, 4 and 8 and the autotuner may pick another.
constexpr int EC_G = 4;

// Lane K's x in every aligned group of G lanes.  Groups of 2 and 4 lanes
// take a DPP quad permutation (one VALU move, which the compiler may fold
// into the instruction that uses it); wider groups a bpermute through the
// LDS crossbar.  Every lane of a group must be active.
template <int G, int K>
__device__ __forceinline__ int grp_bcast(int x) {
  static_assert(K >= 0 && K < G, "source lane inside the group");
  if constexpr (G == 4) {
    return __builtin_amdgcn_mov_dpp(x, K | (K << 2) | (K << 4) | (K << 6), 0xF, 0xF, false);
  } else if constexpr (G == 2) {
    return __builtin_amdgcn_mov_dpp(x, K | (K << 2) | ((K + 2) << 4) | ((K + 2) << 6), 0xF, 0xF, false);
  } else {
    return __shfl(x, K, G);
  }
}
template <int G, int K>
__device__ __forceinline__ float grp_bcast(float x) {
  return __int_as_float(grp_bcast<G, K>(__float_as_int(x)));
}
// lane l's x ^ mask partner (mask < G)
template <int G, int MASK>
__device__ __forceinline__ int grp_xor(int x) {
  if constexpr (G <= 4) {
    constexpr int a = 0 ^ MASK, b = 1 ^ MASK, c = 2 ^ MASK, d = 3 ^ MASK;
    return __builtin_amdgcn_mov_dpp(x, a | (b << 2) | (c << 4) | (d << 6), 0xF, 0xF, false);
  } else {
    return __shfl_xor(x, MASK, G);
  }
}
template <int G, int O = 1>
__device__ __forceinline__ int ec_group_sum(int x) {
  if constexpr (O < G) return ec_group_sum<G, 2 * O>(x + grp_xor<G, O>(x));
  else return x;
}
// acc += lane K's terms, K = 0 .. G-1 in order
template <int G, int K = 0>
__device__ __forceinline__ void ec_ordered_add(float& el, float& Qa, float tx, float ty, float tz, float tq) {
  if constexpr (K < G) {
    el += grp_bcast<G, K>(tx);
    el += grp_bcast<G, K>(ty);
    el += grp_bcast<G, K>(tz);
    Qa += grp_bcast<G, K>(tq);
    ec_ordered_add<G, K + 1>(el, Qa, tx, ty, tz, tq);
  }
}
// the number of lanes K of the group whose iv equals this lane's iu
template <int G, int K = 0>
__device__ __forceinline__ int ec_matches(int iv, int iu) {
  if constexpr (K < G) return (int)(grp_bcast<G, K>(iv) == iu) + ec_matches<G, K + 1>(iv, iu);
  else return 0;
}

// The moved face's term (cusimp_free.cu:497-575), or -1 when it turns over
// (a term lies in [0, 1]).  Two steps are cheaper than the reference's and
// give the same bits:
//  * the edge that does not touch the moved corner keeps its length, read
//    from the slot record (the same expression over the same positions);
//  * the flip test is decided from the unnormalised normal where
//    flip_filter.h proves the answer, else by the reference's
//    normalise-then-dot.
__device__ __forceinline__ float ec_moved_term(const float* __restrict__ P, int4 t, float4 L, float4 op, int moved,
                                               V3 m) {
  const bool mx = t.x == moved, my = t.y == moved, mz = t.z == moved;
  const V3 n0 = mx ? m : ldp(P, t.x);
  const V3 n1 = my ? m : ldp(P, t.y);
  const V3 n2 = mz ? m : ldp(P, t.z);
  const V3 nn = vcross(vsub(n1, n0), vsub(n2, n0));
  int fl = flipf::flip_decide(op.x, op.y, op.z, nn.x, nn.y, nn.z);
  if (fl < 0) fl = vdot(V3{op.x, op.y, op.z}, vdivs(nn, vnorm(nn))) < 0;
  if (fl) return -1.0f;
  const V3 d01 = vsub(n0, n1), d12 = vsub(n1, n2), d20 = vsub(n2, n0);
  float e01, e12, e20;
  if ((int)mx + (int)my + (int)mz == 1) {
    // the two edges at the moved corner; the third from the slot record
    const float lx = vnorm(mz ? d12 : d01);
    const float ly = vnorm(my ? d12 : d20);
    e01 = mz ? L.x : lx;
    e12 = mx ? L.y : (my ? ly : lx);
    e20 = my ? L.z : ly;
  } else {
    e01 = vnorm(d01); e12 = vnorm(d12); e20 = vnorm(d20);
  }
  const float sh = (e01 + e12 + e20) / 2;
  const float area = sqrtf(sh * (sh - e01) * (sh - e12) * (sh - e20));
  const float k43 = 4.0f * 1.73205077648162841796875f;   // 4.0f * sqrtf(3.0f)
  const float term = k43 * area / (e01 * e01) + e12 * e12 + e20 * e20 + 0.0000001f;
  return 1.0f - clamp_ref(term, 0.0f, 1.0f);
}

// Both fans' slots for the edge as one list (fan u, then fan v), G at a
// time: serial length and skinny sums in that order, counts, flip.  A chunk
// may hold the end of fan u and the start of fan v, so an edge takes
// ceil((deg u + deg v) / G) passes instead of one per fan chunk.  Lanes past
// the list's end contribute +0 terms, added like the others: the sums start
// at +0 and hold no -0, so x + (+0) leaves them unchanged.
template <int G>
__device__ __forceinline__ void ec_fans(const float* __restrict__ P, const int4* __restrict__ sface,
                                        const float4* __restrict__ splane, const float4* __restrict__ slen, int ua,
                                        int ub, int va, int vb, int g, int u, int v, V3 m, float& el, float& Qa,
                                        int& ne, int& nt, int& flip) {
  const int nu = ub - ua, n = nu + (vb - va);
  for (int base = 0; base < n; base += G) {
    const int j = base + g;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f, tq = 0.0f;
    if (j < n) {
      const bool fu = j < nu;
      const int s = fu ? ua + j : va + (j - nu);
      const int moved = fu ? u : v, other = fu ? v : u;
      int4 t = sface[s];
      float4 L = slen[s];
      if (t.x > t.y) { tx = L.x; ne++; }
      if (t.y > t.z) { ty = L.y; ne++; }
      if (t.z > t.x) { tz = L.z; ne++; }
      if (!(t.x == other || t.y == other || t.z == other)) {
        tq = ec_moved_term(P, t, L, splane[s], moved, m);
        if (tq < 0.0f) { flip = 1; tq = 0.0f; }
        nt++;
      }
    }
    ec_ordered_add<G>(el, Qa, tx, ty, tz, tq);
  }
}

// the cost's destinations: the mesh's persistent region (when the
// reference writes it) and the claim's per-vertex minima of
// (cost << 32 | edge)
__device__ __forceinline__ void ec_finish(uint32_t* c, u64* vmin, int e, int u, int v, uint32_t val,
                                          bool write_region) {
  if (write_region) *c = val;
  const u64 key = ((u64)val << 32) | (u64)(uint32_t)e;
  atomicMin(&vmin[u], key);
  atomicMin(&vmin[v], key);
}

template <int G>
__global__ void __launch_bounds__(BLOCK) k_edge_cost_g(const float* __restrict__ P, const int4* __restrict__ vedge,
                                                       const int* __restrict__ E_ptr, const int* __restrict__ fan_off,
                                                       const int* __restrict__ fsucc, const int4* __restrict__ sface,
                                                       const float4* __restrict__ splane,
                                                       const float4* __restrict__ slen, const float* __restrict__ Q,
                                                       const uint8_t* __restrict__ table, uint32_t* __restrict__ cost,
                                                       const int* __restrict__ vmesh,
                                                       const long long* __restrict__ voff,
                                                       const long long* __restrict__ eoffm,
                                                       const MeshParam* __restrict__ mp, u64* vmin, int placement,
                                                       const int* __restrict__ gate) {
  if (gated(gate)) return;
  // every lane of a group has the same i, so each return below is taken by
  // the whole group and the width-G shuffles always see all its lanes
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int g = threadIdx.x & (G - 1);
  if (i >= *E_ptr) return;
  int4 ev = vedge[i];
  const int e = ev.x, u = ev.y, v = ev.z;
  int k = vmesh[u];
  const long long pmo = mp[k].pmo;
  long long lu = u - voff[k], lv = v - voff[k];
  uint32_t* c = cost + mp[k].eco + (e - eoffm[k]);
  // where the reference leaves the cost unchanged, the kept value is used
  if (table[pmo + lu] && table[pmo + lv]) {
    if (g == 0) ec_finish(c, vmin, e, u, v, 0xFFFFFFFFu, true);
    return;
  }
  if (mp[k].stuck) {
    if (g == 0) ec_finish(c, vmin, e, u, v, *c, false);
    return;
  }
  int ua = fan_off[u], ub = fan_off[u + 1], va = fan_off[v], vb = fan_off[v + 1];
  // link condition: equal successor pairs; more than 2 -> maximum, not 2 ->
  // unchanged.  The group loads G slots of each fan at once and compares
  // every pair through lane broadcasts (successors are vertex indices, >= 0;
  // lanes past a fan's end hold -1 and -2, which match nothing).
  int dup = 0;
  for (int a0 = ua; a0 < ub; a0 += G) {
    const int iu = a0 + g < ub ? fsucc[a0 + g] : -1;
    for (int b0 = va; b0 < vb; b0 += G) {
      const int iv = b0 + g < vb ? fsucc[b0 + g] : -2;
      dup += ec_matches<G>(iv, iu);
    }
  }
  dup = ec_group_sum<G>(dup);
  if (dup > 2) {
    if (g == 0) ec_finish(c, vmin, e, u, v, 0xFFFFFFFFu, true);
    return;
  }
  if (dup != 2) {
    if (g == 0) ec_finish(c, vmin, e, u, v, *c, false);
    return;
  }
  const float scale = mp[k].scale, tau = mp[k].tau;
  V3 v0 = ldp(P, u), v1 = ldp(P, v);
  V3 m = edge_position(Q, u, v, v0, v1, placement);
  float el = 0.0f, Qa = 0.0f;
  int ne = 0, nt = 0, flip = 0;
  ec_fans<G>(P, sface, splane, slen, ua, ub, va, vb, g, u, v, m, el, Qa, ne, nt, flip);
  ne = ec_group_sum<G>(ne);
  nt = ec_group_sum<G>(nt);
  flip = ec_group_sum<G>(flip);
  if (g != 0) return;
  if (flip) { ec_finish(c, vmin, e, u, v, 0xFFFFFFFFu, true); return; }
  el = el / (float)ne / scale * tau;
  el = el + vnorm(vsub(v0, v1)) / scale * tau;
  Qa = 5.0f * Qa;
  const float* qu = Q + (long long)10 * u;
  const float* qv = Q + (long long)10 * v;
  float q[10];
#pragma unroll
  for (int j = 0; j < 10; j++) q[j] = qu[j] + qv[j];
  float t0 = q[0] * m.x + q[1] * m.y + q[2] * m.z + q[3] * 1.0f;
  float t1 = q[1] * m.x + q[4] * m.y + q[5] * m.z + q[6] * 1.0f;
  float t2 = q[2] * m.x + q[5] * m.y + q[7] * m.z + q[8] * 1.0f;
  float t3 = q[3] * m.x + q[6] * m.y + q[8] * m.z + q[9] * 1.0f;
  float qe = t0 * m.x + t1 * m.y + t2 * m.z + t3 * 1.0f;
  float cc = qe / (scale * scale);
  ec_finish(c, vmin, e, u, v,
            u32sat(clamp_ref(cc + el + Qa / (float)nt * tau, 0.0f, COST_RANGE) / COST_RANGE * U32F), true);
}

// ------------------------------------------------ grouped fan walks
// Kernels that walk the fans of one vertex or one collapse give each item a
// group of FAN_G lanes: the group splits the fan slots and combines its
// lanes' answers with shuffles, so the chain of dependent loads is
// deg / FAN_G long instead of deg.
constexpr int FAN_G = 8;
__device__ __forceinline__ int fan_group_or(int x) {
#pragma unroll
  for (int o = 1; o < FAN_G; o <<= 1) x |= __shfl_xor(x, o, FAN_G);
  return x;
}

// ---------------------------------------------------------------- claim
// The reference takes, per alive face, the minimum key (cost << 32 | edge)
// over every edge whose endpoint's fan holds the face (cusimp_free.cu:
// 587-610).  A face is in the fan of x exactly when x is one of its
// vertices, so that minimum is the minimum over the face's three vertices
// of the per-vertex minimum over incident edges (vmin, two atomics per edge
// in the cost kernel; min is exact in any order).
//
// Winner decision, one lane per vertex: an edge wins when it passes the
// threshold (cusimp_free.cu:650) and holds the minimum key of every face in
// both fans (:658-681).  u and v have non-empty fans (the edge's own face),
// so vmin[u] == vmin[v] == key is necessary: the only candidate of a vertex
// is the edge its minimum names, taken at its first endpoint.  Given that, a
// face holds key as its minimum exactly when no corner of it (the fan slots'
// round-start sface) has a smaller vertex minimum.  win[] was cleared by the
// edge emit.  Winners' fans are disjoint, so applying them later, in any
// order, gives the reference's result.
__global__ void __launch_bounds__(BLOCK) k_winners_v(const int* __restrict__ V_ptr, const u64* __restrict__ vmin,
                                                     const int* __restrict__ edges, const int* __restrict__ vmesh,
                                                     const MeshParam* __restrict__ mp,
                                                     const int* __restrict__ fan_off,
                                                     const int4* __restrict__ sface, int* __restrict__ win,
                                                     const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int V = *V_ptr;
  // a group of FAN_G lanes per vertex; every early return depends on u
  // alone, so it is taken by the whole group
  const int u = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / FAN_G);
  const int g = threadIdx.x & (FAN_G - 1);
  if (u >= V) return;
  const u64 key = vmin[u];
  if (key == ~0ull) return;   // no incident edge
  const int e = (int)(uint32_t)key;
  if (edges[2 * e] != u) return;
  const int v = edges[2 * e + 1];
  if (vmin[v] != key || (uint32_t)(key >> 32) > mp[vmesh[u]].ct) return;
  // every corner of every face of both fans has its minimum at key or above
  const int a0 = fan_off[u], n0 = fan_off[u + 1] - a0, a1 = fan_off[v], n = n0 + fan_off[v + 1] - a1;
  int bad = 0;
  for (int j = g; j < n; j += FAN_G) {
    const int4 t = sface[j < n0 ? a0 + j : a1 + (j - n0)];
    bad |= (int)(vmin[t.x] < key) | (int)(vmin[t.y] < key) | (int)(vmin[t.z] < key);
  }
  if (!fan_group_or(bad) && g == 0) win[e] = 1;
}
// The winners from the dense list (ascending edge order), one group of
// CL_G lanes each: the entry in the mesh's collapsed region, then the
// collapse with the skinny restore (cusimp_free.cu:683-774).  The faces of
// a winner's fans belong to no other winner and still hold their round-start
// corners (sface), so the group's lanes take the fan slots in parallel:
//   surgery    faces with both u and v are removed (by the lane of their
//              fan-u slot only), the other fan-v faces get u for v;
//   skinny     any remaining face of either fan skinny with u at the
//              midpoint (an OR over the lanes: the reference stops at the
//              first, and only "any" is used);
//   restore    positions, occupancy and every fan face from the snapshots.
// Each face is written by one lane, in program order.
constexpr int CL_G = 8;
__device__ void collapse_one(int i, int g, const int* __restrict__ wlist, const int* __restrict__ winpos,
                             const int* __restrict__ edges, const int* __restrict__ vmesh,
                             const long long* __restrict__ eoffm, int* __restrict__ coll,
                             const int* __restrict__ fan_off, const int4* __restrict__ sface, float* P,
                             const float* __restrict__ P0, int* T, const int* __restrict__ T0, int* occ,
                             const float* __restrict__ Q, int placement) {
  const int e = wlist[i];
  const int u = edges[2 * e], v = edges[2 * e + 1];
  const V3 pu = ldp(P, u), pv = ldp(P, v);
  const V3 m = edge_position(Q, u, v, pu, pv, placement);
  const int ua = fan_off[u], ub = fan_off[u + 1], va = fan_off[v], vb = fan_off[v + 1];
  if (g == 0) {
    int k = vmesh[u];
    long long base = eoffm[k];
    coll[2 * base + (i - winpos[base])] = e;
    stp(P, u, m);
    stp(P, v, V3{0.0f, 0.0f, 0.0f});
    occ[v] = 0;
  }
  int sk = 0;
  const int du = ub - ua, n = du + (vb - va);
  for (int j = g; j < n; j += CL_G) {
    const bool in_u = j < du;
    const int s = in_u ? ua + j : va + (j - du);
    const int4 t = sface[s];
    const bool has_u = t.x == u || t.y == u || t.z == u;
    const bool has_v = t.x == v || t.y == v || t.z == v;
    if (has_u && has_v) {   // removed; written once, from its fan-u slot
      if (in_u) { T[3 * t.w] = -1; T[3 * t.w + 1] = -1; T[3 * t.w + 2] = -1; }
      continue;
    }
    I3 c = {t.x, t.y, t.z};
    if (!in_u) {            // fan-v face without u: v becomes u
      if (c.i == v) { c.i = u; T[3 * t.w] = u; }
      else if (c.j == v) { c.j = u; T[3 * t.w + 1] = u; }
      else { c.k = u; T[3 * t.w + 2] = u; }
    }
    const V3 a = c.i == u ? m : ldp(P, c.i), b = c.j == u ? m : ldp(P, c.j), d = c.k == u ? m : ldp(P, c.k);
    sk |= is_skinny(a, b, d) ? 1 : 0;
  }
#pragma unroll
  for (int o = 1; o < CL_G; o <<= 1) sk |= __shfl_xor(sk, o, CL_G);
  if (!sk) return;
  if (g == 0) {
    occ[v] = 1;
    stp(P, u, ldp(P0, u));
    stp(P, v, ldp(P0, v));
  }
  for (int j = g; j < n; j += CL_G) {
    const bool in_u = j < du;
    const int s = in_u ? ua + j : va + (j - du);
    const int4 t = sface[s];
    if (!in_u && (t.x == u || t.y == u || t.z == u)) continue;   // restored from its fan-u slot
    const int f = t.w;
    T[3 * f] = T0[3 * f]; T[3 * f + 1] = T0[3 * f + 1]; T[3 * f + 2] = T0[3 * f + 2];
  }
}
__global__ void __launch_bounds__(BLOCK) k_collapse_b(const int* __restrict__ wlist, const int* __restrict__ W_ptr,
                                                      const int* __restrict__ winpos, const int* __restrict__ edges,
                                                      const int* __restrict__ vmesh,
                                                      const long long* __restrict__ eoffm, int* __restrict__ coll,
                                                      const int* __restrict__ fan_off,
                                                      const int4* __restrict__ sface, float* P,
                                                      const float* __restrict__ P0, int* T,
                                                      const int* __restrict__ T0, int* occ,
                                                      const float* __restrict__ Q, int placement,
                                                      const int* __restrict__ gate) {
  if (gated(gate)) return;
  // grid-stride over winners; every lane of a group has the same i, so the
  // loop exits, returns and shuffles are group-wide
  const int g = threadIdx.x & (CL_G - 1);
  const int W = *W_ptr;
  const int stride = gridDim.x * (BLOCK / CL_G);
  for (int i = (blockIdx.x * blockDim.x + threadIdx.x) / CL_G; i < W; i += stride)
    collapse_one(i, g, wlist, winpos, edges, vmesh, eoffm, coll, fan_off, sface, P, P0, T, T0, occ, Q, placement);
}

// per-mesh edge offsets (faces emit their edges in order)

// After the collapse, in one launch: blocks below nbf remove invalid faces
// (cusimp_free.cu:218-232), the rest count zero-length edges per mesh (the
// line-edge pass runs only for meshes that have one).  The two read and
// write disjoint data.
// The face workgroups also do the first check's initialisation
// (k_check_init's work: nothing between here and the check reads those
// arrays); nbf covers max(Fl, M) threads.
__global__ void k_invalid_line_b(int* T, const int* __restrict__ F_ptr, int nbf, const float* __restrict__ P,
                                 const int* __restrict__ edges, const int* __restrict__ E_ptr,
                                 const int* __restrict__ vmesh, int* lcnt, long long Fl,
                                 int M, uint8_t* __restrict__ qflag, uint8_t* __restrict__ S,
                                 int* __restrict__ seg_a, int* __restrict__ seg_b, int* __restrict__ kparent,
                                 int* __restrict__ nmark, int* __restrict__ novf,
                                 unsigned long long* __restrict__ npairs, int* __restrict__ oflag,
                                 const int* __restrict__ gate) {
  if (gated(gate)) return;
  if ((int)blockIdx.x < nbf) {
    const int F = *F_ptr;
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < Fl) {
      qflag[f] = 0;
      S[f] = 0;
      oflag[f] = 0;
    }
    if (f < M) {
      seg_a[f] = -1;
      seg_b[f] = -1;
    }
    if (f == 0) {
      kparent[0] = -1;
      *nmark = 0;
      *novf = 0;
      *npairs = 0;
    }
    if (f >= F) return;
    I3 t = ldt(T, f);
    if (dead3(t)) return;
    if (t.i == t.j || t.j == t.k || t.k == t.i) { T[3 * f] = -1; T[3 * f + 1] = -1; T[3 * f + 2] = -1; }
    return;
  }
  int e = (blockIdx.x - nbf) * blockDim.x + threadIdx.x;
  if (e >= *E_ptr) return;
  int u = edges[2 * e];
  V3 p = ldp(P, u), q = ldp(P, edges[2 * e + 1]);
  if (p.x == q.x && p.y == q.y && p.z == q.z) atomicAdd(&lcnt[vmesh[u]], 1);
}
// exclusive scan of a small per-mesh int array (x * mult + extra *
// extra_mult) into long long offsets, off[M] = total; all threads of one
// block take part (M may exceed the block)
__device__ void block_small_scan(const int* __restrict__ x, int M, long long* __restrict__ off, long long mult,
                                 const int* __restrict__ extra, long long extra_mult) {
  __shared__ long long carry;
  __shared__ long long part[BLOCK];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < M; base += BLOCK) {
    int i = base + threadIdx.x;
    long long v = 0;
    if (i < M) v = (long long)x[i] * mult + (extra ? (long long)extra[i] * extra_mult : 0);
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < BLOCK; o <<= 1) {
      long long y = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < M) off[i] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == BLOCK - 1) carry += part[BLOCK - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) off[M] = carry;
  __syncthreads();
}
// After the collapse, per mesh, in one block: the winner count; line-edge
// collapse, sequential per mesh (D5), appending to the mesh's collapsed
// region; the scan cut
// (cusimp_free.cu:1016); the flat offsets of the cut entries and of the undo
// outputs; and which meshes take part in the check.  Every check of a round
// (the first and one per undo pass) adds 2 undo entries per marked cut entry,
// so a mesh's undo region holds the incoming list of a stuck mesh plus
// 2 (passes + 1) entries per cut entry, passes being the mode's cap.
__global__ void __launch_bounds__(BLOCK) k_line_cut_tail(int M, const int* __restrict__ lcnt, float* P, int* T,
                                                         int* occ, const int* __restrict__ edges,
                                                         const long long* __restrict__ eoffm,
                                                         const int* __restrict__ fan_off, const int* __restrict__ fan,
                                                         int* coll, const int* __restrict__ winpos,
                                                         int* __restrict__ h, int* __restrict__ ncoll,
                                                         int* __restrict__ cut, long long* __restrict__ cutoff,
                                                         long long* __restrict__ uobase,
                                                         const int* __restrict__ nustk, int* __restrict__ part,
                                                         int strict, const int* __restrict__ gate) {
  if (gated(gate)) return;
  for (int k = threadIdx.x; k < M; k += BLOCK) {
    const int hk = winpos[eoffm[k + 1]] - winpos[eoffm[k]];   // the mesh's winners
    h[k] = hk;
    int n = hk;
    long long base = 2 * eoffm[k];
    if (lcnt[k] > 0) {
      for (long long e = eoffm[k]; e < eoffm[k + 1]; e++) {
        int u = edges[2 * e], v = edges[2 * e + 1];
        V3 p = ldp(P, u), q = ldp(P, v);
        if (!(p.x == q.x && p.y == q.y && p.z == q.z)) continue;
        coll[base + n++] = (int)e;
        stp(P, v, V3{0.0f, 0.0f, 0.0f});
        occ[v] = 0;
        surgery(T, fan_off, fan, u, v);
      }
    }
    ncoll[k] = n;
    // the strict guard scans every collapsed entry
    int ck = strict ? n : (hk == 0 ? 0 : min(n, ((hk + 511) / 512) * 512));
    cut[k] = ck;
    part[k] = ck > 0 ? 1 : 0;
  }
  __syncthreads();
  block_small_scan(cut, M, cutoff, 1, nullptr, 0);
  block_small_scan(cut, M, uobase, undo_slots_per_cut(strict), nustk, 1);
}

// ------------------------------------------------- flat views over meshes
// faces of the fans of the scanned collapsed entries (flat list over meshes)
__global__ void k_qflag_b(const long long* __restrict__ cutoff, int M, const long long* __restrict__ eoffm,
                          const int* __restrict__ coll, const int* __restrict__ edges,
                          const int* __restrict__ fan_off, const int* __restrict__ fan,
                          const int* __restrict__ part, uint8_t* __restrict__ qflag, const int* __restrict__ gate) {
  if (gated(gate)) return;
  // a group of FAN_G lanes per collapse takes the two fans' slots in turn
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / FAN_G;
  const int g = threadIdx.x & (FAN_G - 1);
  if (i >= cutoff[M]) return;
  int k = find_seg(cutoff, M, i);
  if (!part[k]) return;
  int e = coll[2 * eoffm[k] + (i - cutoff[k])];
  const int u = edges[2 * e], v = edges[2 * e + 1];
  const int a0 = fan_off[u], n0 = fan_off[u + 1] - a0, a1 = fan_off[v], n = n0 + fan_off[v + 1] - a1;
  for (int j = g; j < n; j += FAN_G) qflag[fan[j < n0 ? a0 + j : a1 + (j - n0)]] = 1;
}
__global__ void k_query_flags_b(const uint8_t* __restrict__ qflag, const int* __restrict__ T,
                                const int* __restrict__ F_ptr, const int* __restrict__ fmesh,
                                const int* __restrict__ part, int* __restrict__ qf, int* __restrict__ lf) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  int F = *F_ptr;
  if (f > F) return;
  if (f == F) { qf[F] = 0; lf[F] = 0; return; }
  bool alive = T[3 * f] != -1 || T[3 * f + 1] != -1 || T[3 * f + 2] != -1;
  bool p = alive && part[fmesh[f]];
  lf[f] = p ? 1 : 0;
  qf[f] = (p && qflag[f]) ? 1 : 0;
}

// -------------------------------------------------------- segmented LBVH
// leaf boxes of the alive faces of participating meshes
__global__ void __launch_bounds__(BLOCK) k_leaf_boxes_b(const int* __restrict__ leaf_face, const int* __restrict__ n_ptr,
                                                        const float* __restrict__ P, const int* __restrict__ T,
                                                        float4* __restrict__ lbox_lo, float4* __restrict__ lbox_hi) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= *n_ptr) return;
  const int f = leaf_face[l];
  I3 t = ldt(T, f);
  V3 a = ldp(P, t.i), b = ldp(P, t.j), c = ldp(P, t.k);
  // .w of the low corner: the leaf's face (read by the node boxes)
  lbox_lo[LBOX_STRIDE * l] = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)),
                                         fminf(a.z, fminf(b.z, c.z)), __int_as_float(f));
  lbox_hi[LBOX_STRIDE * l] = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)),
                                         fmaxf(a.z, fmaxf(b.z, c.z)), 0.0f);
}
// vertex bounds of one mesh at admission (ordered-int atomics, 6 slots)
__global__ void __launch_bounds__(BLOCK) k_mesh_bounds(const float* __restrict__ P, long long n, int* box) {
  __shared__ float red[6][BLOCK / 64];
  long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float x[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (v < n) {
    x[0] = x[3] = P[3 * v]; x[1] = x[4] = P[3 * v + 1]; x[2] = x[5] = P[3 * v + 2];
  }
#pragma unroll
  for (int q = 0; q < 6; q++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      float y = __shfl_xor(x[q], o, 64);
      x[q] = q < 3 ? fminf(x[q], y) : fmaxf(x[q], y);
    }
  }
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; q++) red[q][threadIdx.x >> 6] = x[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < 6; q++) {
      float y = red[q][0];
      for (int w = 1; w < BLOCK / 64; w++) y = q < 3 ? fminf(y, red[q][w]) : fmaxf(y, red[q][w]);
      if (q < 3) atomicMin(&box[q], f2ord(y));
      else atomicMax(&box[q], f2ord(y));
    }
  }
}
__global__ void k_scene_init_b(int* scene, int M) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * M) return;
  scene[i] = (i % 6) < 3 ? 0x7FFFFFFF : (int)0x80000000;
}
// 64-bit keys: mesh index above the 30-bit Morton code of the box centre in
// the mesh's admission box
// mesh id above the (3 x bits)-bit Morton code of the box centre in the
// mesh's box (bits <= 10 per axis; ties are broken by index in the Karras
// build, so the bit width changes the tree's shape, never the candidates)
__device__ __forceinline__ u64 mesh_morton_key(int k, const MeshParam* __restrict__ mp, float4 lo, float4 hi,
                                               int bits) {
  const float* box = mp[k].box;
  float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  const float cells = (float)(1 << bits);
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float ext = box[3 + a] - box[a];
    float t = ext > 0.0f ? (c[a] - box[a]) / ext : 0.0f;
    t = fminf(fmaxf(t * cells, 0.0f), cells - 1.0f);
    q[a] = (uint32_t)t;
  }
  return ((u64)(uint32_t)k << (3 * bits)) | (u64)(expand10(q[0]) * 4 + expand10(q[1]) * 2 + expand10(q[2]));
}
__global__ void k_morton_b(const float4* __restrict__ lbox_lo, const float4* __restrict__ lbox_hi,
                           const int* __restrict__ n_ptr, const int* __restrict__ leaf_face,
                           const int* __restrict__ fmesh, const MeshParam* __restrict__ mp, u64* __restrict__ keys,
                           int* __restrict__ vals, int bits) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= *n_ptr) return;
  keys[l] = mesh_morton_key(fmesh[leaf_face[l]], mp, lbox_lo[LBOX_STRIDE * l], lbox_hi[LBOX_STRIDE * l], bits);
  vals[l] = l;
}
// the face's box (the expressions of k_leaf_boxes_b)
__device__ __forceinline__ void face_box(const float* __restrict__ P, I3 t, float4& lo, float4& hi) {
  V3 a = ldp(P, t.i), b = ldp(P, t.j), c = ldp(P, t.k);
  lo = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)), 0.0f);
  hi = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)), 0.0f);
}

// ------------------------------------------------- spatial face order
// The exhaustive search walks each mesh's faces in a spatial (Morton) order,
// so the faces of a tile group, and the queries of a wave, are neighbours
// and most group tests miss.  The order (sorder: every face slot once,
// mesh-major) is sorted from the round-start snapshot the first time a
// check needs it, and then carried through each round's face compaction,
// so it stays spatial without further sorts.  Any order gives the same
// candidate set; the order only decides how much the culling saves.
__global__ void k_sorder_keys(const int* __restrict__ T0, const float* __restrict__ P0, int F,
                              const int* __restrict__ fmesh, const MeshParam* __restrict__ mp, u64* __restrict__ keys,
                              int* __restrict__ vals, int bits) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const I3 t = ldt(T0, f);   // every face of the snapshot is alive
  float4 lo, hi;
  face_box(P0, t, lo, hi);
  keys[f] = mesh_morton_key(fmesh[f], mp, lo, hi, bits);
  vals[f] = f;
}
// An admitted mesh's vertex bounds: the ordered-integer min / max of
// k_mesh_bounds decoded to floats (all zero for a mesh without vertices).
__global__ void k_box_store(const int* __restrict__ scene, int any, float* __restrict__ box) {
  const int q = threadIdx.x;
  if (q >= 6) return;
  int i = scene[q];
  i = i >= 0 ? i : i ^ 0x7FFFFFFF;
  box[q] = any ? __int_as_float(i) : 0.0f;
}
// each active mesh's box from its slot's admission bounds
__global__ void k_mp_boxes(MeshParam* __restrict__ mp, int M, const int* __restrict__ kslot,
                           const float* __restrict__ dbox) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const float* b = dbox + 6LL * kslot[k];
#pragma unroll
  for (int q = 0; q < 6; q++) mp[k].box[q] = b[q];
}
// keys of an admitted mesh's faces [f0, f0 + n): the Morton code of the box
// centre in the mesh's vertex box (the mesh is appended last, so its
// faces follow every earlier mesh's in the order)
__global__ void k_sorder_keys_new(const int* __restrict__ T, const float* __restrict__ P, long long f0, int n,
                                  const float* __restrict__ box, u64* __restrict__ keys, int* __restrict__ vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 lo, hi;
  face_box(P, ldt(T, (int)(f0 + i)), lo, hi);
  const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  const float bl[3] = {box[0], box[1], box[2]}, bh[3] = {box[3], box[4], box[5]};
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float ext = bh[a] - bl[a];
    float t = ext > 0.0f ? (c[a] - bl[a]) / ext : 0.0f;
    t = fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
    q[a] = (uint32_t)t;
  }
  keys[i] = (u64)(expand10(q[0]) * 4 + expand10(q[1]) * 2 + expand10(q[2]));
  vals[i] = (int)(f0 + i);
}
// per position of the order: the face's box and index if it takes part in
// the check (alive, in a participating mesh), else an empty box and -1;
// and the union of the boxes of each group of 16 positions and of each tile
// of BLOCK positions (one workgroup), aligned to the position grid and
// unbounded if a box has a NaN coordinate
__global__ void __launch_bounds__(BLOCK) k_sorder_boxes(const int* __restrict__ sorder, const int* __restrict__ F_ptr,
                                                        const int* __restrict__ fmesh, const int* __restrict__ part,
                                                        const float* __restrict__ P,
                                                        const int* __restrict__ T, int* __restrict__ sid,
                                                        float4* __restrict__ slo, float4* __restrict__ shi,
                                                        float4* __restrict__ tlo, float4* __restrict__ thi,
                                                        float4* __restrict__ glo, float4* __restrict__ ghi,
                                                        const int* __restrict__ gate) {
  __shared__ float4 wlo[BLOCK / 64], whi[BLOCK / 64];
  if (gated(gate)) return;   // uniform over the workgroup
  const int F = *F_ptr;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  if (p < F) {
    const int f = sorder[p];
    const I3 tf = ldt(T, f);
    const bool in = !dead3(tf) && part[fmesh[f]];   // alive, in a participating mesh
    if (in) face_box(P, tf, lo, hi);
    sid[p] = in ? f : -1;
    slo[p] = lo;
    shi[p] = hi;
  }
  float l0 = lo.x, l1 = lo.y, l2 = lo.z, h0 = hi.x, h1 = hi.y, h2 = hi.z;
  if ((l0 != l0) | (l1 != l1) | (l2 != l2) | (h0 != h0) | (h1 != h1) | (h2 != h2))
    l0 = l1 = l2 = -INFINITY, h0 = h1 = h2 = INFINITY;
  for (int o = 1; o < 64; o <<= 1) {
    if (o == 16 && (threadIdx.x & 15) == 0) {   // the union of each group of 16 positions
      glo[p >> 4] = make_float4(l0, l1, l2, 0.0f);
      ghi[p >> 4] = make_float4(h0, h1, h2, 0.0f);
    }
    l0 = fminf(l0, __shfl_xor(l0, o, 64));
    l1 = fminf(l1, __shfl_xor(l1, o, 64));
    l2 = fminf(l2, __shfl_xor(l2, o, 64));
    h0 = fmaxf(h0, __shfl_xor(h0, o, 64));
    h1 = fmaxf(h1, __shfl_xor(h1, o, 64));
    h2 = fmaxf(h2, __shfl_xor(h2, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    wlo[threadIdx.x >> 6] = make_float4(l0, l1, l2, 0.0f);
    whi[threadIdx.x >> 6] = make_float4(h0, h1, h2, 0.0f);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float4 a = wlo[0], c = whi[0];
    for (int w = 1; w < BLOCK / 64; w++) {
      a = make_float4(fminf(a.x, wlo[w].x), fminf(a.y, wlo[w].y), fminf(a.z, wlo[w].z), 0.0f);
      c = make_float4(fmaxf(c.x, whi[w].x), fmaxf(c.y, whi[w].y), fmaxf(c.z, whi[w].z), 0.0f);
    }
    tlo[blockIdx.x] = a;
    thi[blockIdx.x] = c;
  }
}
// queries (flagged faces) listed in the order; qpos per position
// (a query: alive, in a participating mesh, and flagged)
struct SorderQueryOp {
  const int* sorder;
  const int* F_ptr;   // live faces (the order's length)
  const int* T;
  const int* fmesh;
  const int* part;
  const uint8_t* qflag;
  int* qpos;
  int* qlist;
  struct Rec { int f, q; };
  __device__ Rec load(long long p) const {
    const long long F = *F_ptr;
    const int f = sorder[max(min(p, F - 1), 0LL)];
    const I3 t = ldt(T, f);
    Rec r;
    r.f = f;
    r.q = (p < F && !dead3(t) && qflag[f] && part[fmesh[f]]) ? 1 : 0;
    return r;
  }
  __device__ int count(const Rec& r) const { return r.q; }
  __device__ void gather(long long, Rec&) const {}
  __device__ void emit(long long p, const Rec& r, int pos) const {
    qpos[p] = pos;
    if (r.q) qlist[pos] = r.f;
  }
};
// the order through the round's face compaction: kept faces, renamed
struct SorderCompactOp {
  const int* sorder;
  const int* F_ptr;   // live faces (the order's length)
  const int* kpos;
  int* out;
  struct Rec { int keep, to; };
  __device__ Rec load(long long p) const {
    const long long F = *F_ptr;
    const int f = sorder[max(min(p, F - 1), 0LL)];
    const int a = kpos[f], b = kpos[f + 1];
    Rec r;
    r.keep = p < F ? b - a : 0;
    r.to = a;
    return r;
  }
  __device__ int count(const Rec& r) const { return r.keep; }
  __device__ void gather(long long, Rec&) const {}
  __device__ void emit(long long, const Rec& r, int pos) const {
    if (r.keep) out[pos] = r.to;
  }
};

// queries in Morton order: flag per sorted leaf position, then compaction
__global__ void k_query_sorted(const int* __restrict__ sorted_leaf, const int* __restrict__ leaf_face,
                               const int* __restrict__ n_ptr, const uint8_t* __restrict__ qflag, int* __restrict__ qf) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int n = *n_ptr;
  if (l > n) return;
  qf[l] = l < n ? (int)qflag[leaf_face[sorted_leaf[l]]] : 0;
}
__global__ void k_query_list(const int* __restrict__ sorted_leaf, const int* __restrict__ leaf_face,
                             const int* __restrict__ n_ptr, const int* __restrict__ qf, const int* __restrict__ qpos,
                             int* __restrict__ qlist, int* __restrict__ nq) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int n = *n_ptr;
  if (l == n) *nq = qpos[n];
  if (l >= n || !qf[l]) return;
  qlist[qpos[l]] = leaf_face[sorted_leaf[l]];
}

// Stable LSD radix sort of up to SS_MAX (u64 key, int value) pairs in one
// workgroup and one launch, for small alive-face counts where the
// multi-launch sort is all launch latency.  Keys and values stay in LDS
// across passes.  Each pass sorts 4 bits: every thread owns SS_ITEMS
// consecutive items and one private counter per digit (packed two to a
// u32: digit r in the low half, digit r+8 in the high half of row r), a
// block scan over the counters in (row, thread) order gives every item its
// stable position, and the items are written back in place.
constexpr int SS_NT = BLOCK1K, SS_ITEMS = 8, SS_MAX = SS_NT * SS_ITEMS;
__global__ void __launch_bounds__(SS_NT) k_sort_small(const u64* __restrict__ kin, const int* __restrict__ vin,
                                                      const int* __restrict__ n_ptr, int kbits,
                                                      u64* __restrict__ kout, int* __restrict__ vout, int* err) {
  __shared__ u64 sk[SS_MAX];
  __shared__ int sv[SS_MAX];
  __shared__ unsigned cnt[8 * SS_NT];
  __shared__ unsigned wl[SS_NT / 64];
  const int n = *n_ptr;
  const int t = threadIdx.x;
  if (n > SS_MAX) {   // the host launches this only under a bound; never expected
    if (t == 0) atomicAdd(err, 1);
    return;
  }
#pragma unroll
  for (int i = 0; i < SS_ITEMS; i++) {
    int idx = t * SS_ITEMS + i;
    sk[idx] = idx < n ? kin[idx] : ~0ull;   // padding sorts last and stays there
    sv[idx] = idx < n ? vin[idx] : -1;
  }
  __syncthreads();
  for (int shift = 0; shift < kbits; shift += 4) {
    u64 k[SS_ITEMS];
    int v[SS_ITEMS];
    unsigned lr[SS_ITEMS];
#pragma unroll
    for (int r = 0; r < 8; r++) cnt[r * SS_NT + t] = 0;
#pragma unroll
    for (int i = 0; i < SS_ITEMS; i++) {
      k[i] = sk[t * SS_ITEMS + i];
      v[i] = sv[t * SS_ITEMS + i];
      int d = (int)((k[i] >> shift) & 15ull);
      int r = d & 7, h = d >> 3;
      unsigned c = cnt[r * SS_NT + t];
      lr[i] = (c >> (16 * h)) & 0xFFFFu;
      cnt[r * SS_NT + t] = c + (1u << (16 * h));
    }
    __syncthreads();
    // exclusive scan of the 8*SS_NT packed counters in (row, thread) order;
    // thread t owns entries [8t, 8t+8)
    unsigned x[8], pre[8], sum = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) { x[j] = cnt[8 * t + j]; pre[j] = sum; sum += x[j]; }
    unsigned tot;
    unsigned off = block_exclusive_scan_1024(sum, wl, tot);
#pragma unroll
    for (int j = 0; j < 8; j++) cnt[8 * t + j] = off + pre[j];
    __syncthreads();
    const unsigned low_total = tot & 0xFFFFu;   // items with digits 0..7
    int pos[SS_ITEMS];
#pragma unroll
    for (int i = 0; i < SS_ITEMS; i++) {
      int d = (int)((k[i] >> shift) & 15ull);
      int r = d & 7, h = d >> 3;
      unsigned base = (cnt[r * SS_NT + t] >> (16 * h)) & 0xFFFFu;
      pos[i] = (int)(base + lr[i] + (h ? low_total : 0u));
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SS_ITEMS; i++) { sk[pos[i]] = k[i]; sv[pos[i]] = v[i]; }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < SS_ITEMS; i++) {
    int idx = t * SS_ITEMS + i;
    if (idx < n) { kout[idx] = sk[idx]; vout[idx] = sv[idx]; }
  }
}

// stable LSD radix sort of (u64 key, int value), 8-bit digits (as
// k_rs_hist/k_rs_scatter, 64-bit keys)
__global__ void __launch_bounds__(BLOCK) k_rs_hist64(const u64* __restrict__ keys, const int* __restrict__ n_ptr,
                                                     int shift, int ntiles, int* __restrict__ hist) {
  __shared__ int hh[256];
  int n = *n_ptr;
  hh[threadIdx.x] = 0;
  __syncthreads();
  long long base = (long long)blockIdx.x * RS_TILE;
  for (int i = 0; i < RS_ITEMS; i++) {
    long long idx = base + (long long)i * BLOCK + threadIdx.x;
    if (idx < n) atomicAdd(&hh[(int)((keys[idx] >> shift) & 255ull)], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * ntiles + blockIdx.x] = hh[threadIdx.x];
}
// LDS slot of the split's position i: one slot of padding per eight, so the
// readback of eight consecutive 64-bit keys per lane (lanes 64 bytes apart,
// two bank positions) spreads over the banks
__device__ __forceinline__ int rs_skew(int i) { return i + (i >> 3); }
__global__ void __launch_bounds__(BLOCK) k_rs_scatter64(const u64* __restrict__ kin, const int* __restrict__ vin,
                                                        const int* __restrict__ n_ptr, int shift, int ntiles,
                                                        const int* __restrict__ hoff, u64* __restrict__ kout,
                                                        int* __restrict__ vout) {
  __shared__ u64 sk[RS_TILE + RS_TILE / 8];   // slot i at rs_skew(i)
  __shared__ int sv[RS_TILE + RS_TILE / 8];
  __shared__ int lds[BLOCK / 64];
  __shared__ int dstart[256];
  __shared__ int goff[256];
  int n = *n_ptr;
  long long base = (long long)blockIdx.x * RS_TILE;
  if (base >= n) return;
  int cnt = (int)min((long long)RS_TILE, (long long)n - base);
  u64 k[RS_ITEMS];
  int v[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int li = threadIdx.x * RS_ITEMS + i;
    if (li < cnt) { k[i] = kin[base + li]; v[i] = vin[base + li]; }
    else { k[i] = ~0ull; v[i] = -1; }
  }
  goff[threadIdx.x] = hoff[(long long)threadIdx.x * ntiles + blockIdx.x];
  for (int bit = 0; bit < 8; bit++) {
    int ones = 0;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) ones += (int)((k[i] >> (shift + bit)) & 1ull);
    int total_zeros;
    int zbefore = block_exclusive_scan(RS_ITEMS - ones, lds, total_zeros);
    int obefore = threadIdx.x * RS_ITEMS - zbefore;
    int zc = 0, oc = 0;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
      int b = (int)((k[i] >> (shift + bit)) & 1ull);
      int dst = b ? (total_zeros + obefore + oc++) : (zbefore + zc++);
      sk[rs_skew(dst)] = k[i];
      sv[rs_skew(dst)] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
      k[i] = sk[rs_skew(threadIdx.x * RS_ITEMS + i)];
      v[i] = sv[rs_skew(threadIdx.x * RS_ITEMS + i)];
    }
    __syncthreads();
  }
  dstart[threadIdx.x] = RS_TILE;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int li = threadIdx.x * RS_ITEMS + i;
    if (li < cnt) {
      int d = (int)((k[i] >> shift) & 255ull);
      if (li == 0 || (int)((sk[rs_skew(li - 1)] >> shift) & 255ull) != d) dstart[d] = li;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int li = threadIdx.x * RS_ITEMS + i;
    if (li < cnt) {
      int d = (int)((k[i] >> shift) & 255ull);
      long long o = (long long)goff[d] + (li - dstart[d]);
      kout[o] = k[i];
      vout[o] = v[i];
    }
  }
}

__device__ __forceinline__ int kdelta64(const u64* __restrict__ key, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  u64 a = key[i], b = key[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll((long long)(a ^ b));
}
__global__ void k_karras64(const u64* __restrict__ key, const int* __restrict__ n_ptr, int* __restrict__ left,
                           int* __restrict__ right, int* __restrict__ parent, int* __restrict__ rfirst,
                           int* __restrict__ rlast) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = *n_ptr;
  if (i >= n - 1) return;
  int d = (kdelta64(key, n, i, i + 1) - kdelta64(key, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = kdelta64(key, n, i, i - d);
  int lmax = 2;
  while (kdelta64(key, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta64(key, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = kdelta64(key, n, i, j);
  int s = 0;
  for (int div = 2;; div *= 2) {
    int t = (l + div - 1) / div;
    if (kdelta64(key, n, i, i + (s + t) * d) > dnode) s += t;
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
// sorted leaf ranges per mesh: seg_a[k] / seg_b[k] (first / last sorted
// leaf); the mesh id sits above the code_bits-bit Morton code of the key
__global__ void k_seg_bounds(const u64* __restrict__ key, const int* __restrict__ n_ptr, int* __restrict__ seg_a,
                             int* __restrict__ seg_b, int code_bits) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  int n = *n_ptr;
  if (l >= n) return;
  int k = (int)(key[l] >> code_bits);
  if (l == 0 || (int)(key[l - 1] >> code_bits) != k) seg_a[k] = l;
  if (l == n - 1 || (int)(key[l + 1] >> code_bits) != k) seg_b[k] = l;
}
// subtree root of mesh k (the node whose range is exactly [a, b]: a radix
// tree has one, and a Karras node is numbered by one end of its range) and
// the node after that subtree in depth-first order
__global__ void k_mesh_roots(int M, const int* __restrict__ n_ptr, const int* __restrict__ seg_a,
                             const int* __restrict__ seg_b, const int* __restrict__ rfirst,
                             const int* __restrict__ rlast, int* __restrict__ root, int* __restrict__ root_end) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  int n = *n_ptr;
  int a = seg_a[k], b = seg_b[k];
  if (a < 0) { root[k] = -1; root_end[k] = -1; return; }
  int r;
  if (a == b) r = n - 1 + a;
  else if (a <= n - 2 && rfirst[a] == a && rlast[a] == b) r = a;
  else r = b;
  root[k] = r;
  if (b == n - 1) root_end[k] = -1;
  else root_end[k] = (b + 1 <= n - 2 && rfirst[b + 1] == b + 1) ? (b + 1) : (n - 1 + b + 1);
}

// traversal from the query's mesh root to the end of that subtree
template <int STRICT>
__global__ void __launch_bounds__(BLOCK) k_traverse_b(const int* __restrict__ qlist, const int* __restrict__ nq_ptr,
                                                      const int* __restrict__ n_ptr, const int* __restrict__ fmesh,
                                                      const int* __restrict__ root, const int* __restrict__ root_end,
                                                      const float4* __restrict__ nodeA, const float4* __restrict__ nodeB,
                                                      const float* __restrict__ P, const int* __restrict__ T,
                                                      uint8_t* __restrict__ S, unsigned long long* __restrict__ stats,
                                                      int* err) {
  int qi = blockIdx.x * blockDim.x + threadIdx.x;
  int nq = *nq_ptr;
  unsigned cand = 0;
  bool hit = false;
  if (qi < nq) {
    const long long cap = 2LL * (2LL * (*n_ptr) - 1) + 8;
    int f = qlist[qi];
    int k = fmesh[f];
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
    for (int a = 0; a < 3; a++) {
      qlo[a] = fminf(pf[0][a], fminf(pf[1][a], pf[2][a]));
      qhi[a] = fmaxf(pf[0][a], fmaxf(pf[1][a], pf[2][a]));
    }
    int node = root[k], end = root_end[k];
    long long steps = 0;
    while (node != end && node >= 0) {
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
          if (pair_hits_g<STRICT>(pf, pg, tf, tg)) { hit = true; break; }
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

// ------------------------------------------------ traversal in two passes
// The walk and the narrow phase run in separate kernels.  In one kernel a
// wave runs the long FP64 pair test in every loop step in which any of its
// lanes sits on an overlapping leaf, so its time is (steps of the longest
// walk) x (pair test) whatever the number of queries.  Here the walk only
// collects, per query, the overlapping faces (a chain of node loads and
// nothing else; pair_hits rejects the pairs that share an edge), and the
// pair kernel tests one pair per lane.  The verdict per query is the OR over
// its pairs, the same as the walk that stops at the first hit.
//
// A query may be split over 2^kshift lanes, each walking one subtree at
// depth kshift of the mesh's tree: the subtrees partition the leaves, so the
// union of the lanes' candidates is the query's candidate set, and the
// longest dependent chain of node loads shrinks for queries with many
// candidates (what bounds the kernel when there are few queries).
//
// Candidates collect in LDS (slot j of lane t at j*BLOCK+t, conflict-free);
// a lane whose list fills moves it to the pair buffer with an atomic of its
// own, and after the walk each wave reserves room for what its lanes still
// hold with one atomic.  Pairs are (query face, other face).  A query whose
// pairs do not fit the pair buffer goes to the overflow list, which
// k_traverse_b walks in full.
constexpr int CAND_CAP = 32;
constexpr int N_XCD = 8;   // MI355X: 8 accelerator complex dies, 32 CUs and one L2 each
__global__ void __launch_bounds__(BLOCK) k_traverse_emit(
    const int* __restrict__ qlist, const int* __restrict__ nq_ptr, const int* __restrict__ n_ptr,
    const int* __restrict__ fmesh, const int* __restrict__ root, const int* __restrict__ root_end,
    const float4* __restrict__ nodeA, const float4* __restrict__ nodeB, const float* __restrict__ P,
    const int* __restrict__ T, int2* __restrict__ pairs, long long pair_cap, unsigned long long* __restrict__ npairs,
    int* __restrict__ ovf, int* __restrict__ novf, unsigned long long* __restrict__ stats, int* err, int kshift) {
  __shared__ int cl[CAND_CAP * BLOCK];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const long long gid = (long long)blockIdx.x * blockDim.x + t;
  {
    // Warm each XCD's L2 with the node arrays before the walks (a kernel
    // starts with a cold L2, and a walk is a chain of dependent node loads).
    // Workgroups go to the XCDs round-robin; the ones on one XCD split the
    // nodes between them.  Only for trees that fit an L2 comfortably.
    const long long nn = 2LL * (*n_ptr) - 1;
    if (nn > 0 && nn * 32 <= (2LL << 20)) {
      const long long per = (gridDim.x + N_XCD - 1) / N_XCD;           // workgroups per XCD
      const long long j = blockIdx.x / N_XCD;                            // index within the XCD
      const long long chunk = (nn + per - 1) / per;
      const long long a = j * chunk, b = min(nn, a + chunk);
      float acc = 0.0f;
      // one load per 128-byte line (8 nodes of 16 bytes)
      for (long long i = a + 8LL * t; i < b; i += 8LL * BLOCK) acc += nodeA[i].x + nodeB[i].x;
      if (acc == 1.0e-37f && stats) atomicAdd(&stats[3], 1ull);   // keeps the loads; practically never true
    }
  }
  const long long qi = gid >> kshift;
  const int part = (int)(gid & ((1 << kshift) - 1));
  int nq = *nq_ptr;
  int cnt = 0;
  int f = -1;
  bool full = false;   // the pair buffer ran out: walk this query in full instead
  if (qi < nq) {
    const long long cap = 2LL * (2LL * (*n_ptr) - 1) + 8;
    f = qlist[qi];
    int k = fmesh[f];
    I3 tf = ldt(T, f);
    float qlo[3], qhi[3];
    {
      V3 a = ldp(P, tf.i), b = ldp(P, tf.j), c = ldp(P, tf.k);
      qlo[0] = fminf(a.x, fminf(b.x, c.x)); qhi[0] = fmaxf(a.x, fmaxf(b.x, c.x));
      qlo[1] = fminf(a.y, fminf(b.y, c.y)); qhi[1] = fmaxf(a.y, fmaxf(b.y, c.y));
      qlo[2] = fminf(a.z, fminf(b.z, c.z)); qhi[2] = fmaxf(a.z, fmaxf(b.z, c.z));
    }
    // this lane's subtree: 2^kshift parts of the mesh's tree at depth
    // kshift (the right child of a node is its left child's escape); a
    // leaf reached early belongs to the part whose remaining bits are zero
    int node = root[k], end = root_end[k];
    bool mine_part = true;
    for (int b = kshift - 1; b >= 0 && mine_part; b--) {
      int c = __float_as_int(nodeB[node].w);
      if (c < 0) {
        mine_part = (part & ((2 << b) - 1)) == 0;
        break;
      }
      node = ((part >> b) & 1) ? __float_as_int(nodeA[c].w) : c;
    }
    if (kshift > 0) end = __float_as_int(nodeA[node].w);
    if (!mine_part) node = end;
    long long steps = 0;
    while (node != end && node >= 0) {
      if (++steps > cap) { atomicAdd(err, 1); break; }
      float4 A = nodeA[node], B = nodeB[node];
      const bool ov = !((qhi[0] < A.x) | (B.x < qlo[0]) | (qhi[1] < A.y) | (B.y < qlo[1]) | (qhi[2] < A.z) |
                        (B.z < qlo[2]));   // no short circuits: the comparisons are pure
      int esc = __float_as_int(A.w);
      if (ov) {
        int c = __float_as_int(B.w);
        if (c >= 0) { node = c; continue; }
        int g = ~c;
        if (g != f) {
          if (cnt == CAND_CAP) {   // LDS list full: move it to the pair buffer
            long long b = (long long)atomicAdd(npairs, (unsigned long long)CAND_CAP);
            for (int j = 0; j < CAND_CAP; j++)
              if (b + j < pair_cap) pairs[b + j] = make_int2(f, cl[j * BLOCK + t]);
            cnt = 0;
            if (b + CAND_CAP > pair_cap) { full = true; break; }
          }
          cl[cnt * BLOCK + t] = g;
          cnt++;
        }
      }
      node = esc;
    }
  }
  // wave reservation for what is left in LDS
  const bool small = f >= 0 && !full;
  int mine = small ? cnt : 0;
  int incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  int total = __shfl(incl, 63, 64);
  unsigned long long base = 0;
  if (lane == 63 && total) base = atomicAdd(npairs, (unsigned long long)total);
  base = __shfl(base, 63, 64);
  long long at = (long long)base + (incl - mine);
  // a list that crosses the end of the buffer still fills the slots below
  // it (every slot under min(npairs, pair_cap) is written by one lane) and
  // its query is also walked in full, whose verdict then stands
  const bool fits = small && at + mine <= pair_cap;
  if (f >= 0 && !fits) ovf[atomicAdd(novf, 1)] = f;
  for (int j = 0; j < mine; j++)
    if (at + j < pair_cap) pairs[at + j] = make_int2(f, cl[j * BLOCK + t]);
}

// One candidate pair per lane (grid-stride over the pairs the walk wrote):
// both orientations of the reference narrow phase; a hit flags the query.
// pairs handled by blocks bid of nb (grid-stride)
template <int STRICT>
__device__ void pair_test_blocks(long long bid, long long nb, const int2* __restrict__ pairs,
                                 const unsigned long long* __restrict__ npairs, long long pair_cap,
                                 const float* __restrict__ P, const int* __restrict__ T, uint8_t* __restrict__ S,
                                 unsigned long long* __restrict__ stats) {
  const unsigned long long n_found = *npairs;   // every candidate the search reserved
  long long n = (long long)n_found;
  if (n > pair_cap) n = pair_cap;
  // the check's candidate and pair counts, once per check (the searches do
  // not count: one address updated by every workgroup serialises them)
  if (stats && bid == 0 && threadIdx.x == 0) {
    atomicAdd(&stats[0], n_found);
    atomicAdd(&stats[2], (unsigned long long)n);
  }
  unsigned hits = 0;
  for (long long i = bid * blockDim.x + threadIdx.x; i < n; i += nb * blockDim.x) {
    int2 pr = pairs[i];
    I3 tf = ldt(T, pr.x), tg = ldt(T, pr.y);
    float pf[3][3], pg[3][3];
    {
      V3 a = ldp(P, tf.i), b = ldp(P, tf.j), c = ldp(P, tf.k);
      pf[0][0] = a.x; pf[0][1] = a.y; pf[0][2] = a.z;
      pf[1][0] = b.x; pf[1][1] = b.y; pf[1][2] = b.z;
      pf[2][0] = c.x; pf[2][1] = c.y; pf[2][2] = c.z;
    }
    {
      V3 a = ldp(P, tg.i), b = ldp(P, tg.j), c = ldp(P, tg.k);
      pg[0][0] = a.x; pg[0][1] = a.y; pg[0][2] = a.z;
      pg[1][0] = b.x; pg[1][1] = b.y; pg[1][2] = b.z;
      pg[2][0] = c.x; pg[2][1] = c.y; pg[2][2] = c.z;
    }
    if (pair_hits_g<STRICT>(pf, pg, tf, tg)) {
      S[pr.x] = 1;
      hits++;
    }
  }
  if (stats) {
    unsigned long long h64 = hits;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) h64 += __shfl_xor(h64, o, 64);
    if ((threadIdx.x & 63) == 0 && h64) atomicAdd(&stats[1], h64);
  }
}
template <int STRICT>
__global__ void __launch_bounds__(BLOCK) k_pair_test(const int2* __restrict__ pairs,
                                                     const unsigned long long* __restrict__ npairs, long long pair_cap,
                                                     const float* __restrict__ P, const int* __restrict__ T,
                                                     uint8_t* __restrict__ S, unsigned long long* __restrict__ stats) {
  pair_test_blocks<STRICT>(blockIdx.x, gridDim.x, pairs, npairs, pair_cap, P, T, S, stats);
}

// --------------------------------------------- exhaustive candidate search
// For small meshes the tree costs more than it saves: a check then tests
// every query against every alive face of its mesh.  The candidate set is
// the same (alive faces g != f whose float AABB overlaps the query's,
// touching included, with the boxes of k_leaf_boxes_b), so the flags are
// the same.  Faces and queries are taken in the spatial order (positions
// foff[k] .. foff[k+1] hold mesh k; dead faces have empty boxes and id -1).
//
// Work unit: one wave of BF_NT queries of one mesh (a query wave, bx)
// against one tile of BF_TILE positions, tiles aligned to the global
// position grid (a tile at a mesh boundary takes only the mesh's
// positions).  The culling is exact at every level, since every union
// contains its members' boxes and the overlap test never separates a NaN
// coordinate (a union with one is unbounded):
//   - k_bf_waves: each query wave stores its queries' boxes and their
//     union, tests the union against the union of every tile of its mesh
//     (from k_sorder_boxes), and lists the (wave, tile) pairs that meet;
//   - k_bf_emit: persistent workgroups take listed pairs; the tile's boxes
//     go to LDS, groups of BF_GRP positions whose union (also from
//     k_sorder_boxes) misses the wave's union are skipped, the rest are
//     tested box by box.
// The spatial order keeps a wave's queries and a tile's faces local, so
// most pairs are never listed, and no workgroup is spent on them.
// Candidates go through per-lane LDS lists, wave-wide flushes and the pair
// buffer.
constexpr int BF_TILE = 256;
constexpr int BF_GRP = 16;
constexpr int BF_NT = 64;   // queries (threads) per workgroup: one wave
static_assert(BF_TILE == BLOCK, "tile unions come from one k_sorder_boxes workgroup");
static_assert(BF_TILE % BF_NT == 0 && BF_NT % BF_GRP == 0 && BF_NT == 64, "one wave, tile load by groups");
// union of the box over `width` consecutive lanes (width a power of two <= 64)
__device__ __forceinline__ void lane_union(float& l0, float& l1, float& l2, float& h0, float& h1, float& h2,
                                           int width) {
  for (int o = 1; o < width; o <<= 1) {
    l0 = fminf(l0, __shfl_xor(l0, o, 64));
    l1 = fminf(l1, __shfl_xor(l1, o, 64));
    l2 = fminf(l2, __shfl_xor(l2, o, 64));
    h0 = fmaxf(h0, __shfl_xor(h0, o, 64));
    h1 = fmaxf(h1, __shfl_xor(h1, o, 64));
    h2 = fmaxf(h2, __shfl_xor(h2, o, 64));
  }
}
__device__ __forceinline__ float wave_uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__global__ void __launch_bounds__(BLOCK) k_bf_blocks(int M, const long long* __restrict__ foff,
                                                     const int* __restrict__ qpos, int* __restrict__ nb,
                                                     long long* __restrict__ boff, int* __restrict__ npair,
                                                     const int* __restrict__ gate) {
  if (gated(gate)) return;
  for (int k = threadIdx.x; k < M; k += BLOCK) {
    int nq = qpos[foff[k + 1]] - qpos[foff[k]];
    nb[k] = (nq + BF_NT - 1) / BF_NT;
  }
  if (threadIdx.x == 0) *npair = 0;   // the (wave, tile) list of this check
  __syncthreads();
  block_small_scan(nb, M, boff, 1, nullptr, 0);   // query waves -> wave offsets per mesh
}
// One workgroup per query wave: its queries' boxes, their union, and the
// listed tiles of its mesh whose union meets it.
__global__ void __launch_bounds__(BF_NT) k_bf_waves(
    const int* __restrict__ qlist, const int* __restrict__ qpos, const long long* __restrict__ foff,
    const long long* __restrict__ boff, int M, const float* __restrict__ P, const int* __restrict__ T,
    const float4* __restrict__ tlo, const float4* __restrict__ thi, float4* __restrict__ qlo,
    float4* __restrict__ qhi, float4* __restrict__ wlo, float4* __restrict__ whi, int4* __restrict__ wrec,
    int2* __restrict__ plist, int* __restrict__ npair, const int* __restrict__ gate) {
  if (gated(gate)) return;
  const int t = threadIdx.x;   // == lane
  const long long bx = blockIdx.x;
  if (bx >= boff[M]) return;   // workgroup-uniform exits
  const int k = find_seg(boff, M, bx);
  const long long a0 = foff[k], a1 = foff[k + 1];
  const long long q1 = qpos[a1];
  const long long qi = qpos[a0] + (bx - boff[k]) * BF_NT + t;
  float l0 = INFINITY, l1 = INFINITY, l2 = INFINITY, h0 = -INFINITY, h1 = -INFINITY, h2 = -INFINITY;
  if (qi < q1) {
    I3 tf = ldt(T, qlist[qi]);
    V3 a = ldp(P, tf.i), b = ldp(P, tf.j), c = ldp(P, tf.k);
    l0 = fminf(a.x, fminf(b.x, c.x)), h0 = fmaxf(a.x, fmaxf(b.x, c.x));
    l1 = fminf(a.y, fminf(b.y, c.y)), h1 = fmaxf(a.y, fmaxf(b.y, c.y));
    l2 = fminf(a.z, fminf(b.z, c.z)), h2 = fmaxf(a.z, fmaxf(b.z, c.z));
    qlo[qi] = make_float4(l0, l1, l2, 0.f);
    qhi[qi] = make_float4(h0, h1, h2, 0.f);
  }
  if ((l0 != l0) | (l1 != l1) | (l2 != l2) | (h0 != h0) | (h1 != h1) | (h2 != h2))
    l0 = l1 = l2 = -INFINITY, h0 = h1 = h2 = INFINITY;
  lane_union(l0, l1, l2, h0, h1, h2, 64);
  const float w0 = wave_uniform(l0), w1 = wave_uniform(l1), w2 = wave_uniform(l2);
  const float x0 = wave_uniform(h0), x1 = wave_uniform(h1), x2 = wave_uniform(h2);
  if (t == 0) {
    wlo[bx] = make_float4(w0, w1, w2, 0.f);
    whi[bx] = make_float4(x0, x1, x2, 0.f);
    const long long q0 = qi;   // lane 0's query: the wave's first
    wrec[bx] = make_int4((int)q0, (int)min(q0 + BF_NT, q1), (int)a0, (int)a1);   // queries, mesh positions
  }
  const long long tt0 = a0 / BF_TILE, tt1 = (a1 - 1) / BF_TILE;   // a1 > a0: the mesh has queries
  for (long long c0 = tt0; c0 <= tt1; c0 += BF_NT) {
    const long long tt = c0 + t;
    bool ov = false;
    if (tt <= tt1) {
      const float4 ul = tlo[tt], uh = thi[tt];
      ov = !((x0 < ul.x) | (uh.x < w0) | (x1 < ul.y) | (uh.y < w1) | (x2 < ul.z) | (uh.z < w2));
    }
    const unsigned long long m = __ballot(ov);
    if (!m) continue;
    int base = 0;
    if (t == 0) base = atomicAdd(npair, __popcll(m));
    base = __shfl(base, 0, 64);
    if (ov) plist[base + __popcll(m & ((1ull << t) - 1))] = make_int2((int)bx, (int)tt);
  }
}
// A listed pair's tile may be split over WAVES waves: each wave takes every
// WAVES-th group of the tile for the same BF_NT queries (lane = query), with
// its own candidate lists, so a pair's serial chain of group tests is WAVES
// times shorter.  Small checks list few pairs and wait on their longest
// chain (one wave per pair: 40 us for a 20k-face check, 13-17 us with four
// waves); large checks list many and are bound by throughput, where one wave
// per pair pays each pair's fixed costs (tile load, reservation) once.  The
// candidate set is the same either way: each group is tested by one wave.
constexpr int BF_WAVES_SMALL = 4;
constexpr long long BF_SPLIT_MAX_FACES = 65536;   // alive faces of a check that take the split
template <int WAVES> struct BfCap { static constexpr int value = WAVES == 1 ? 32 : 16; };
static_assert((BF_TILE / BF_GRP) % BF_WAVES_SMALL == 0, "groups split evenly over the waves");
template <int WAVES>
__global__ void __launch_bounds__(BF_NT * WAVES) k_bf_emit(
    const int2* __restrict__ plist, const int* __restrict__ npair, const int4* __restrict__ wrec,
    const float4* __restrict__ wlo, const float4* __restrict__ whi, const float4* __restrict__ qlo,
    const float4* __restrict__ qhi, const int* __restrict__ qlist, int F, const int* __restrict__ sid,
    const float4* __restrict__ slo, const float4* __restrict__ shi, const float4* __restrict__ gulo,
    const float4* __restrict__ guhi, int2* __restrict__ pairs, long long pair_cap,
    unsigned long long* __restrict__ npairs, int* __restrict__ ovf, int* __restrict__ novf,
    int* __restrict__ oflag, const int* __restrict__ gate) {
  constexpr int BF_CAP = BfCap<WAVES>::value;   // candidate slots per lane and wave before a flush
  __shared__ int cls[WAVES][BF_CAP * BF_NT];
  __shared__ float4 blo[BF_TILE], bhi[BF_TILE];
  __shared__ int bid[BF_TILE];
  __shared__ float4 glo[BF_TILE / BF_GRP], ghi[BF_TILE / BF_GRP];
  const int tid = threadIdx.x;
  const int t = tid & (BF_NT - 1);   // lane == query of the wave
  const int wv = tid / BF_NT;        // this wave's share of the tile's groups
  int* cl = cls[wv];
  const int np = gated(gate) ? 0 : *npair;   // uniform over the grid
  for (int it = blockIdx.x; it < np; it += gridDim.x) {   // workgroup-uniform loop
    // the pair, then the wave's record and the tile together (the tile's
    // loads wait only for the pair; the mesh range masks them afterwards)
    const int2 pe = plist[it];
    const long long bx = pe.x, tile0 = (long long)pe.y * BF_TILE;
    const int4 wr = wrec[bx];
    const long long a0 = wr.z, a1 = wr.w;
    __syncthreads();   // the previous pair's boxes are no longer read
    {
      // the tile's boxes, all loads issued before the LDS stores
      constexpr int R = BF_TILE / (BF_NT * WAVES);
      float4 tl[R], th[R];
      int ti[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const long long pc = min(tile0 + r * BF_NT * WAVES + tid, (long long)F - 1);
        tl[r] = slo[pc];
        th[r] = shi[pc];
        ti[r] = sid[pc];
      }
      // group unions (precomputed over whole groups; at a mesh boundary they
      // also cover the other mesh's boxes, which only weakens the culling)
      if (tid < BF_TILE / BF_GRP) {
        glo[tid] = gulo[tile0 / BF_GRP + tid];
        ghi[tid] = guhi[tile0 / BF_GRP + tid];
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int j = r * BF_NT * WAVES + tid;
        const long long p = tile0 + j;
        const bool in = p >= a0 && p < a1;   // other meshes' positions: never a candidate
        blo[j] = in ? tl[r] : make_float4(INFINITY, INFINITY, INFINITY, 0.f);
        bhi[j] = in ? th[r] : make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
        bid[j] = in ? ti[r] : -1;
      }
    }
    __syncthreads();
    const long long qi = (long long)wr.x + t;
    const long long q1 = wr.y;
    int cnt = 0;
    int f = -1;
    bool full = false;
    float qlo0 = INFINITY, qlo1 = INFINITY, qlo2 = INFINITY, qhi0 = -INFINITY, qhi1 = -INFINITY, qhi2 = -INFINITY;
    if (qi < q1) {
      f = qlist[qi];
      const float4 ql = qlo[qi], qh = qhi[qi];
      qlo0 = ql.x, qlo1 = ql.y, qlo2 = ql.z, qhi0 = qh.x, qhi1 = qh.y, qhi2 = qh.z;
    }
    const float4 wl = wlo[bx], wh = whi[bx];
    const float w0 = wave_uniform(wl.x), w1 = wave_uniform(wl.y), w2 = wave_uniform(wl.z);
    const float x0 = wave_uniform(wh.x), x1 = wave_uniform(wh.y), x2 = wave_uniform(wh.z);
    // The six comparisons of the overlap test, evaluated without short
    // circuits (they have no side effects) so the box loads are not
    // serialised behind branches; four boxes per step.  The append is
    // predicated: every box writes the lane's next list slot and only a hit
    // advances the count, so the step has no branch.  Every lane runs the
    // loop (lanes without a query hold an empty box and never hit), so the
    // flushes are wave-wide: when any lane's list could overrun, the wave
    // moves all lists to the pair buffer with one prefix sum and one
    // atomic.  If the buffer is full the wave stops, and all its queries
    // take the overflow walk.
    for (int gi = wv; gi < BF_TILE / BF_GRP && !full; gi += WAVES) {
      const float4 gl = glo[gi], gh = ghi[gi];
      if ((x0 < gl.x) | (gh.x < w0) | (x1 < gl.y) | (gh.y < w1) | (x2 < gl.z) | (gh.z < w2)) continue;
      // the group is skipped when no lane's own box meets its union (lanes
      // without a query hold an empty box); a wave's queries can be spread
      // over the mesh, so this culls more than the wave's union
      if (!__any(!((qhi0 < gl.x) | (gh.x < qlo0) | (qhi1 < gl.y) | (gh.y < qlo1) | (qhi2 < gl.z) |
                   (gh.z < qlo2))))
        continue;
      for (int j0 = gi * BF_GRP; j0 < gi * BF_GRP + BF_GRP; j0 += 4) {
        if (__any(cnt > BF_CAP - 4)) {
          int incl = cnt;
#pragma unroll
          for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(incl, o, 64);
            if (t >= o) incl += y;
          }
          const int wtot = __shfl(incl, 63, 64);
          unsigned long long b0 = 0;
          if (t == 0) b0 = atomicAdd(npairs, (unsigned long long)wtot);
          b0 = ((unsigned long long)(unsigned)__shfl((int)(b0 >> 32), 0, 64) << 32) |
               (unsigned long long)(unsigned)__shfl((int)(unsigned)b0, 0, 64);
          const long long at = (long long)b0 + (incl - cnt);
          for (int q = 0; q < cnt; q++)
            if (at + q < pair_cap) pairs[at + q] = make_int2(f, cl[q * BF_NT + t]);
          cnt = 0;
          if ((long long)b0 + wtot > pair_cap) { full = true; break; }
        }
        unsigned sep[4];
        int gg[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          float4 lo = blo[j0 + u], hi = bhi[j0 + u];
          gg[u] = bid[j0 + u];
          sep[u] = (unsigned)(qhi0 < lo.x) | (unsigned)(hi.x < qlo0) | (unsigned)(qhi1 < lo.y) |
                   (unsigned)(hi.y < qlo1) | (unsigned)(qhi2 < lo.z) | (unsigned)(hi.z < qlo2);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int g = gg[u];
          const int hit = (int)(sep[u] == 0u) & (int)(g >= 0) & (int)(g != f) & (int)(f >= 0);
          cl[cnt * BF_NT + t] = g;
          cnt += hit;
        }
      }
    }
    // the wave's final reservation for this pair: one prefix sum, one atomic
    const bool small = f >= 0 && !full;
    const int mine = small ? cnt : 0;
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(incl, o, 64);
      if (t >= o) incl += y;
    }
    const int wtot = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (t == 0 && wtot) base = atomicAdd(npairs, (unsigned long long)wtot);
    base = ((unsigned long long)(unsigned)__shfl((int)(base >> 32), 0, 64) << 32) |
           (unsigned long long)(unsigned)__shfl((int)(unsigned)base, 0, 64);
    const long long at = (long long)base + (incl - mine);
    const bool fits = small && at + mine <= pair_cap;
    // a query may overflow in several tiles; it enters the list once
    if (f >= 0 && !fits && atomicExch(&oflag[f], 1) == 0) ovf[atomicAdd(novf, 1)] = f;
    for (int j = 0; j < mine; j++)
      if (at + j < pair_cap) pairs[at + j] = make_int2(f, cl[j * BF_NT + t]);
  }
}
// overflow queries of the exhaustive search: every alive face of the mesh,
// with the narrow phase inline, S[f] = 1 on the first hit
struct BfFull {
  const int* ovf;
  const int* novf;
  const int* fmesh;
  const long long* foff;
  const int* sid;        // by position of the spatial order (-1: not a candidate)
  const float4* slo;
  const float4* shi;
};
template <int STRICT>
__device__ void bf_full_query(long long qi, const BfFull& a, const float* __restrict__ P, const int* __restrict__ T,
                              uint8_t* __restrict__ S) {
  const int* __restrict__ ovf = a.ovf;
  const int* __restrict__ fmesh = a.fmesh;
  const long long* __restrict__ foff = a.foff;
  const int* __restrict__ sid = a.sid;
  const float4* __restrict__ slo = a.slo;
  const float4* __restrict__ shi = a.shi;
  if (qi >= *a.novf) return;
  int f = ovf[qi];
  int k = fmesh[f];
  long long a0 = foff[k], a1 = foff[k + 1];
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
  for (int x = 0; x < 3; x++) {
    qlo[x] = fminf(pf[0][x], fminf(pf[1][x], pf[2][x]));
    qhi[x] = fmaxf(pf[0][x], fmaxf(pf[1][x], pf[2][x]));
  }
  bool hit = false;
  for (long long l = a0; l < a1 && !hit; l++) {
    float4 lo = slo[l], hi = shi[l];
    if (qhi[0] < lo.x || hi.x < qlo[0] || qhi[1] < lo.y || hi.y < qlo[1] || qhi[2] < lo.z || hi.z < qlo[2]) continue;
    int g = sid[l];
    if (g < 0 || g == f) continue;
    I3 tg = ldt(T, g);
    float pg[3][3];
    V3 a = ldp(P, tg.i), b = ldp(P, tg.j), c = ldp(P, tg.k);
    pg[0][0] = a.x; pg[0][1] = a.y; pg[0][2] = a.z;
    pg[1][0] = b.x; pg[1][1] = b.y; pg[1][2] = b.z;
    pg[2][0] = c.x; pg[2][1] = c.y; pg[2][2] = c.z;
    hit = pair_hits_g<STRICT>(pf, pg, tf, tg);
  }
  if (hit) S[f] = 1;
}
// the flat pair test (blocks below nbp) and the overflow queries (the rest)
// of an exhaustive check, in one launch: both only set flags
template <int STRICT>
__global__ void __launch_bounds__(BLOCK) k_pairs_bf(int nbp, const int2* __restrict__ pairs,
                                                    const unsigned long long* __restrict__ npairs, long long pair_cap,
                                                    BfFull a, const float* __restrict__ P, const int* __restrict__ T,
                                                    uint8_t* __restrict__ S, unsigned long long* __restrict__ stats,
                                                    const int* __restrict__ gate) {
  if (gated(gate)) return;
  if ((int)blockIdx.x < nbp) pair_test_blocks<STRICT>(blockIdx.x, nbp, pairs, npairs, pair_cap, P, T, S, stats);
  else bf_full_query<STRICT>((long long)(blockIdx.x - nbp) * blockDim.x + threadIdx.x, a, P, T, S);
}

// --------------------------------------------------------------- marks
__global__ void k_mark_b(const long long* __restrict__ cutoff, int M, const long long* __restrict__ eoffm,
                         const int* __restrict__ coll, const int* __restrict__ edges,
                         const int* __restrict__ fan_off, const int* __restrict__ fan, const int* __restrict__ part,
                         const uint8_t* __restrict__ S, int* __restrict__ mflag, int* __restrict__ total,
                         const int* __restrict__ gate) {
  if (gated(gate)) return;
  // a group of FAN_G lanes per collapse: hit if a face of either fan is flagged
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long i = gid / FAN_G;
  const int g = threadIdx.x & (FAN_G - 1);
  long long tot = cutoff[M];
  if (gid == 0) *total = (int)tot;   // live length of the scan of the marks
  if (i > tot) return;
  if (i == tot) {
    if (g == 0) mflag[tot] = 0;
    return;
  }
  int k = find_seg(cutoff, M, i);
  int hit = 0;
  if (part[k]) {
    int e = coll[2 * eoffm[k] + (i - cutoff[k])];
    const int u = edges[2 * e], v = edges[2 * e + 1];
    const int a0 = fan_off[u], n0 = fan_off[u + 1] - a0, a1 = fan_off[v], n = n0 + fan_off[v + 1] - a1;
    for (int j = g; j < n; j += FAN_G) hit |= (int)(S[fan[j < n0 ? a0 + j : a1 + (j - n0)]] != 0);
  }
  hit = fan_group_or(hit);
  if (g == 0) mflag[i] = hit;
}
// marked entries -> per-mesh marked list (flat, base cutoff[k]) and endpoint
// pairs (per mesh, base 2 cutoff[k]); blocks from nb0 on count marks per
// mesh: hm[k] = marks of mesh k
__global__ void k_marked_b(const long long* __restrict__ cutoff, int M, const long long* __restrict__ eoffm,
                           const int* __restrict__ coll, const int* __restrict__ edges, const int* __restrict__ mflag,
                           const int* __restrict__ mpos, int* __restrict__ marked, int* __restrict__ tmp, int nb0,
                           int* __restrict__ hm, int* __restrict__ part, int* __restrict__ nvu,
                           int* __restrict__ n_active, int* __restrict__ attn, const int* __restrict__ gate) {
  if (gated(gate)) return;
  if ((int)blockIdx.x >= nb0) {
    int k = (blockIdx.x - nb0) * blockDim.x + threadIdx.x;
    if (k >= M) return;
    int c = part[k] ? (mpos[cutoff[k + 1]] - mpos[cutoff[k]]) : 0;
    hm[k] = c;
    nvu[k] += 2 * c;
    part[k] = c > 0 ? 1 : 0;
    if (c > 0) {
      atomicAdd(n_active, 1);
      *attn = 1;   // the host runs the undo passes before the round's end
    }
    return;
  }
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cutoff[M] || !mflag[i]) return;
  int k = find_seg(cutoff, M, i);
  long long j = mpos[i] - mpos[cutoff[k]];
  int e = coll[2 * eoffm[k] + (i - cutoff[k])];
  marked[cutoff[k] + j] = e;
  tmp[2 * cutoff[k] + 2 * j] = edges[2 * e];
  tmp[2 * cutoff[k] + 2 * j + 1] = edges[2 * e + 1];
}
// One undo pass in one launch: blocks below nbu restore the marked edges of
// participating meshes (cusimp_free.cu:829-867); the rest write this pass's
// undo-list entries through the stale map (cusimp_free.cu:869-878), which
// reads only the endpoint pairs and the previous round's map.
__global__ void k_undo_b(int nbu, const long long* __restrict__ cutoff, int M, const int* __restrict__ hm,
                         const int* __restrict__ marked, const int* __restrict__ edges, float* P,
                         const float* __restrict__ P0, int* T, const int* __restrict__ T0, int* occ,
                         const int* __restrict__ fan_off, const int* __restrict__ fan, const int* __restrict__ tmp,
                         const long long* __restrict__ voff, const MeshParam* __restrict__ mp,
                         const int* __restrict__ pm, const long long* __restrict__ uobase,
                         const int* __restrict__ first, int* __restrict__ undo_out, int* err) {
  if ((int)blockIdx.x < nbu) {
    // a group of FAN_G lanes per marked collapse (restore_edge with the
    // fan slots split over the group; every restorer of a face writes the
    // same snapshot)
    long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / FAN_G;
    const int g = threadIdx.x & (FAN_G - 1);
    if (i >= cutoff[M]) return;
    int k = find_seg(cutoff, M, i);
    long long j = i - cutoff[k];
    if (j >= hm[k]) return;
    int e = marked[cutoff[k] + j];
    const int u = edges[2 * e], v = edges[2 * e + 1];
    if (g == 0) {
      occ[v] = 1;
      stp(P, u, ldp(P0, u));
      stp(P, v, ldp(P0, v));
    }
    const int a0 = fan_off[u], n0 = fan_off[u + 1] - a0, a1 = fan_off[v], n = n0 + fan_off[v + 1] - a1;
    for (int q = g; q < n; q += FAN_G) {
      const int f = fan[q < n0 ? a0 + q : a1 + (q - n0)];
      const I3 t0 = ldt(T0, f);
      T[3 * f] = t0.i; T[3 * f + 1] = t0.j; T[3 * f + 2] = t0.k;
    }
    return;
  }
  long long i = (long long)(blockIdx.x - nbu) * blockDim.x + threadIdx.x;
  if (i >= 2 * cutoff[M]) return;
  int k = find_seg(cutoff, M, i / 2);
  long long j = i - 2 * cutoff[k];
  if (j >= 2LL * hm[k]) return;
  int x = tmp[2 * cutoff[k] + j] - (int)voff[k];   // local vertex index
  int r = 0;
  if (x >= 0 && (long long)x < mp[k].pmcap) r = pm[mp[k].pmo + x];
  else atomicAdd(err, 1);
  const long long o = uobase[k] + first[k] + j;
  if (o < uobase[k + 1]) undo_out[o] = r;   // the region holds every pass of the mode's cap
  else atomicAdd(err, 1);
}
// after an undo pass: the next entry of each mesh's undo list, and the
// passes each mesh took part in (those with marks)
__global__ void k_advance_first(int M, const int* __restrict__ hm, int* __restrict__ first, int* __restrict__ npass) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M || hm[k] == 0) return;
  first[k] += 2 * hm[k];
  npass[k] += 1;
}
// unwritten tail of each undo list (D6) -> 0

// ---------------------------------------------------- end of round
// local vertex maps into the mesh regions (the scan over occupancy is
// global; local = global - value at the mesh start)
// Workgroups from nbv on write the undo-list tails: per mesh, zeros from
// the first entry the undo passes left free up to nvu.
__global__ void k_local_map(const int* __restrict__ g, const long long* __restrict__ voff, int M,
                            const int* __restrict__ vmesh, const MeshParam* __restrict__ mp, const int* __restrict__ V_ptr,
                            int* __restrict__ pm, int nbv, const long long* __restrict__ uobase,
                            const int* __restrict__ first, const int* __restrict__ nvu, int* __restrict__ undo_out,
                            const int* __restrict__ gate) {
  if (gated(gate)) return;   // the round's end waits for the undo passes
  if ((int)blockIdx.x >= nbv) {
    const int k = blockIdx.x - nbv;
    if (k >= M) return;
    const long long n = min((long long)nvu[k], uobase[k + 1] - uobase[k]);
    for (long long j = first[k] + threadIdx.x; j < n; j += blockDim.x) undo_out[uobase[k] + j] = 0;
    return;
  }
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int V = *V_ptr;
  if (v >= V) return;
  int k = vmesh[v];
  pm[mp[k].pmo + (v - voff[k])] = g[v] - g[voff[k]];
  if (v == voff[k + 1] - 1) pm[mp[k].pmo + (v - voff[k]) + 1] = g[v + 1] - g[voff[k]];
}
// new per-mesh offsets and sizes, and the offsets of next round's packed
// undo input (one block)
__global__ void __launch_bounds__(BLOCK) k_new_offsets(int M, const long long* __restrict__ voff,
                                                       const long long* __restrict__ foff, const int* __restrict__ g,
                                                       const int* __restrict__ fpos, long long* __restrict__ nvoff,
                                                       long long* __restrict__ nfoff, const int* __restrict__ nvu,
                                                       long long* __restrict__ uioff_new, const int* __restrict__ gate) {
  if (gated(gate)) return;   // the round's end waits for the undo passes
  for (int k = threadIdx.x; k <= M; k += BLOCK) {
    nvoff[k] = g[voff[k]];
    nfoff[k] = fpos[foff[k]];
  }
  block_small_scan(nvu, M, uioff_new, 1, nullptr, 0);
}
// next round's packed undo input from the per-mesh undo outputs
__global__ void k_pack_undo(int M, const long long* __restrict__ uioff_new, const long long* __restrict__ uobase,
                            const int* __restrict__ undo_out, int* __restrict__ undo_in, const int* __restrict__ gate) {
  if (gated(gate)) return;   // the round's end waits for the undo passes
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= uioff_new[M]) return;
  int k = find_seg(uioff_new, M, i);
  undo_in[i] = undo_out[uobase[k] + (i - uioff_new[k])];
}

// The round's end on the device (simp_cuda/pamo/__init__.py:142-158, per
// mesh): the new sizes and the termination state, the stuck flag and the
// stuck undo count the next round reads, the next round's offsets and
// totals, and the attention flag, set when a mesh finished (the host
// retires it before the next round).  Unless forced it does nothing while
// the attention flag is set: the round's first check marked collapses, and
// the host runs it again, forced, after the undo passes.  One block.
__global__ void __launch_bounds__(BLOCK) k_round_status(int M, const long long* __restrict__ nvoff,
                                                        const long long* __restrict__ nfoff,
                                                        const long long* __restrict__ uioff2, MeshParam* mp,
                                                        MeshState* __restrict__ ms, int* __restrict__ nustk,
                                                        long long* __restrict__ voff, long long* __restrict__ foff,
                                                        long long* __restrict__ uioff, int* live, int force) {
  __shared__ int ndone;
  if (!force && live[LIVE_ATTN]) return;   // uniform over the block
  if (threadIdx.x == 0) ndone = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < M; k += BLOCK) {
    MeshState s = ms[k];
    const long long target = mp[k].target;
    const int nprev = s.F;
    const int ncur = (int)(nfoff[k + 1] - nfoff[k]);
    s.rounds++;
    s.F = ncur;
    s.V = (int)(nvoff[k + 1] - nvoff[k]);
    s.n_undo = (int)(uioff2[k + 1] - uioff2[k]);
    bool done = false;
    if (target >= 0 && (ncur <= target || ncur <= 10)) {
      s.exit = ncur <= target ? 1 : 2;
      done = true;
    } else {
      if (ncur == nprev) s.count++;
      else { s.count = 0; s.stuck = 0; }
      if (s.count >= 2) s.stuck = 1;
      if (s.count == mp[k].tolerance) { s.exit = 3; done = true; }
    }
    if (!done && s.rounds >= mp[k].iters) { s.exit = 4; done = true; }
    s.done = done ? 1 : 0;
    ms[k] = s;
    mp[k].stuck = s.stuck;
    nustk[k] = s.stuck ? s.n_undo : 0;
    if (done) atomicAdd(&ndone, 1);
  }
  for (int k = threadIdx.x; k <= M; k += BLOCK) {
    voff[k] = nvoff[k];
    foff[k] = nfoff[k];
    uioff[k] = uioff2[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    live[LIVE_V] = (int)nvoff[M];
    live[LIVE_F] = (int)nfoff[M];
    live[LIVE_UIN] = (int)uioff2[M];
    live[LIVE_DONE] = ndone;
    live[LIVE_ATTN] = ndone > 0 ? 1 : 0;
    live[LIVE_EPOCH] = (int)(((unsigned)live[LIVE_EPOCH] + SCAN_EPOCH_STEP) & SCAN_EPOCH_MASK);
  }
}

// ---------------------------------------------------------- retirement
__global__ void k_faces_local(const int* __restrict__ T, long long n, int voff, int* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = T[i] - voff;
}
// keep[i] = 1 when element i belongs to a mesh that stays (keep[n] = 0)
__global__ void k_keep_by_mesh(const int* __restrict__ mesh, long long n, const int* __restrict__ newk,
                               int* __restrict__ keep) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  keep[i] = (i < n && newk[mesh[i]] >= 0) ? 1 : 0;
}
__global__ void k_retire_verts(const float* __restrict__ P, const int* __restrict__ vmesh, const int* __restrict__ keep,
                               const int* __restrict__ pos, const int* __restrict__ newk, long long V,
                               float* __restrict__ outP, int* __restrict__ outm) {
  long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V || !keep[v]) return;
  int o = pos[v];
  outP[3LL * o] = P[3 * v]; outP[3LL * o + 1] = P[3 * v + 1]; outP[3LL * o + 2] = P[3 * v + 2];
  outm[o] = newk[vmesh[v]];
}
__global__ void k_retire_faces(const int* __restrict__ T, const int* __restrict__ fmesh, const int* __restrict__ keep,
                               const int* __restrict__ pos, const int* __restrict__ vpos, const int* __restrict__ newk,
                               long long F, int* __restrict__ outT, int* __restrict__ outm) {
  long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F || !keep[f]) return;
  int o = pos[f];
  outT[3LL * o] = vpos[T[3 * f]]; outT[3LL * o + 1] = vpos[T[3 * f + 1]]; outT[3LL * o + 2] = vpos[T[3 * f + 2]];
  outm[o] = newk[fmesh[f]];
}
// undo inputs of kept meshes to their new packed offsets (newuo[k] < 0: dropped)
__global__ void k_retire_undo2(const int* __restrict__ undo_in, const long long* __restrict__ uioff, int M,
                               const long long* __restrict__ newuo, int* __restrict__ out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= uioff[M]) return;
  int k = find_seg(uioff, M, i);
  if (newuo[k] < 0) return;
  out[newuo[k] + (i - uioff[k])] = undo_in[i];
}

// Persistent regions of the kept meshes moved down to their new offsets
// (kept mesh j: source offset src[j], destination dst[j], dst[nk] = total),
// so the regions take space only for active meshes.  grid.y: 0 = the vertex
// maps and invalid-vertex tables (same offsets), 1 = the edge costs.
__global__ void k_move_regions(int nk, const long long* __restrict__ pm_src, const long long* __restrict__ pm_dst,
                               const long long* __restrict__ ec_src, const long long* __restrict__ ec_dst,
                               const int* __restrict__ pm, const uint8_t* __restrict__ table,
                               const uint32_t* __restrict__ ecost, int* __restrict__ pm2, uint8_t* __restrict__ table2,
                               uint32_t* __restrict__ ecost2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.y == 0) {
    if (i >= pm_dst[nk]) return;
    int j = find_seg(pm_dst, nk, i);
    long long from = pm_src[j] + (i - pm_dst[j]);
    pm2[i] = pm[from];
    table2[i] = table[from];
  } else {
    if (i >= ec_dst[nk]) return;
    int j = find_seg(ec_dst, nk, i);
    ecost2[i] = ecost[ec_src[j] + (i - ec_dst[j])];
  }
}

}  // namespace eng
}  // namespace pamo
