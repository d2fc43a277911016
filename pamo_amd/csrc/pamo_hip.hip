// Host orchestration and C ABI of libpamo_hip.so (see include/pamo_hip.h).
//
// One context = one simplifier on one GPU.  Its persistent state mirrors
// the reference's CUSimp_Free (simp_cuda/src/cusimp_free.h:184-280): the
// vertex map, the invalid-vertex table and the edge costs survive between
// rounds and are (re)allocated by the same n + n/5 growth rule, zero-filled
// on (re)allocation.  Everything else is round scratch, sized once for the
// largest mesh and reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdarg>
#include <initializer_list>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>
#include <mutex>

#include "../../include/pamo_hip.h"
#include "pamo_bvh.h"
#include "pamo_engine_kernels.h"
#include "pamo_kernels.h"
#include "remesh_kernels.h"
#include "stage3_kernels.h"

using namespace pamo;

namespace {

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

enum Counter {
  C_E = 0, C_NFAN, C_NWIN, C_NCOLL, C_NLINE, C_CUT, C_NALIVE, C_NQUERY, C_NMARK, C_ERR, C_NEWV, C_NEWF,
  C_V, C_F, C_CUTTOT, C_NOVF, C_COUNT = 16
};

inline unsigned nblk(long long n) { return (unsigned)std::max<long long>(1, (n + BLOCK - 1) / BLOCK); }

// cusimp_free.cu:923: u32(clamp(tau, 0, 10) / 10 * UINT32_MAX), clamp as
// cusimp_free.cu:29-34 (NaN -> upper bound), saturating conversion
inline uint32_t collapse_threshold(float tau) {
  float m = 0.0f > tau ? 0.0f : tau;
  float t = m < COST_RANGE ? m : COST_RANGE;
  float x = t / COST_RANGE * U32F;
  if (!(x > 0.0f)) return 0u;
  if (x >= U32F) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

}  // namespace

struct Engine;

// Launch-shape parameters of the engine.  None changes a result: they pick
// work distributions whose outputs are identical (lanes per edge of the cost
// kernel, Morton bits of the LBVH and spatial-order keys, the face counts at
// which the exhaustive search hands over to the LBVH and splits its pairs
// over waves, the walk's lanes per query, the exhaustive search's persistent
// workgroups).  The defaults are the hand-measured values; pamo_set_tuning
// or a PAMO_TUNING_FILE (a flat JSON object of these names) overrides them.
struct Tuning {
  long long ec_lanes = 4;               // lanes per edge of the cost kernel (2, 4 or 8)
  long long morton_bits = 10;           // Morton bits per axis (6..10)
  long long bf_max_faces = 262144;      // exhaustive search for meshes up to this many faces, else the LBVH
  long long bf_split_max_faces = 65536; // exhaustive pairs split over four waves up to this many alive faces
  long long walk8_max = 65536;          // LBVH walk: 8 lanes per query up to this many faces ...
  long long walk4_max = 262144;         // ... 4 lanes up to this many, else 1
  long long emit_groups = 2048;         // persistent workgroups of the exhaustive search's pair kernel
  long long pipeline = 1;               // enqueue the next round before reading this one's result (0: never)
  long long graph = 1;                  // record rounds into hipGraphs and replay them (one mesh; 0: never) ...
  long long graph_max_faces = 32768;    // ... up to this many faces
  long long s3_bf_max_faces = 16384;    // stage 3: exhaustive contact search up to this many faces, else LBVHs
  long long s3_cg_one_block = 32768;    // stage 3: one-workgroup CG steps up to this many vertices
  long long s3_graph = 1;               // stage 3: record each Newton iteration's tail into a hipGraph (0: never)
};
struct TuningField { const char* name; long long Tuning::*field; long long lo, hi; };
static const TuningField kTuningFields[] = {
    {"ec_lanes", &Tuning::ec_lanes, 2, 8},
    {"morton_bits", &Tuning::morton_bits, 6, 10},
    {"bf_max_faces", &Tuning::bf_max_faces, 0, 1LL << 40},
    {"bf_split_max_faces", &Tuning::bf_split_max_faces, 0, 1LL << 40},
    {"walk8_max", &Tuning::walk8_max, 0, 1LL << 40},
    {"walk4_max", &Tuning::walk4_max, 0, 1LL << 40},
    {"emit_groups", &Tuning::emit_groups, 1, 65536},
    {"pipeline", &Tuning::pipeline, 0, 1},
    {"graph", &Tuning::graph, 0, 1},
    {"graph_max_faces", &Tuning::graph_max_faces, 0, 1LL << 40},
    {"s3_bf_max_faces", &Tuning::s3_bf_max_faces, 0, 1LL << 40},
    {"s3_cg_one_block", &Tuning::s3_cg_one_block, 0, 1LL << 40},
    {"s3_graph", &Tuning::s3_graph, 0, 1},
};
static bool tuning_set(Tuning& t, const char* name, long long v) {
  for (const TuningField& f : kTuningFields)
    if (strcmp(f.name, name) == 0) {
      if (v < f.lo || v > f.hi) return false;
      if (f.field == &Tuning::ec_lanes && v != 2 && v != 4 && v != 8) return false;
      t.*(f.field) = v;
      return true;
    }
  return false;
}
// a flat JSON object {"name": integer, ...}; unknown names are ignored
static void tuning_load_file(Tuning& t, const char* path) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return;
  std::string txt;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), fp)) > 0) txt.append(buf, n);
  fclose(fp);
  size_t i = 0;
  while ((i = txt.find('"', i)) != std::string::npos) {
    size_t j = txt.find('"', i + 1);
    if (j == std::string::npos) break;
    std::string key = txt.substr(i + 1, j - i - 1);
    size_t k = txt.find(':', j);
    if (k == std::string::npos) break;
    char* end = nullptr;
    long long v = strtoll(txt.c_str() + k + 1, &end, 10);
    if (end != txt.c_str() + k + 1) (void)tuning_set(t, key.c_str(), v);
    i = end ? (size_t)(end - txt.c_str()) : k + 1;
  }
}

struct S3State;
struct pamo_ctx {
  Engine* eng = nullptr;   // batched engine state (pamo_engine.inc)
  S3State* s3 = nullptr;   // stage-3 system (pamo_stage3.inc)
  int device = 0;
  int shuffle_mode = PAMO_SHUFFLE_FEISTEL;
  hipStream_t st = nullptr;
  std::string err;

  // persistent state with the reference's growth rule
  size_t alloc_pts = 0, alloc_edge_cost = 0;
  Buf pts_map, table, edge_cost, undo_list, tmp_undo;
  size_t undo_cap = 0, tmp_cap = 0;   // ints

  // round buffers
  Buf P, P0, occ, T, T0, cnt, fan_off, fan, ecnt, eoff, edges, Q, tri_min, win, winpos, collapsed;
  Buf qflag, qf, qpos, af, apos, qlist, leaf_face, lbox_lo, lbox_hi, keys0, keys1, vals0, vals1, hist;
  Buf pairs, ovf, seg_lo, seg_hi;
  Buf kleft, kright, kparent, rfirst, rlast, nodeA, nodeB, S, S_first, mflag, mpos, marked;
  Buf scan_desc, counters, cstats, shuf_flags, shuf_pos;
  unsigned scan_epoch = 0;   // epoch 0 never names a call, so zeroed words read as unpublished
  // Graph capture of a round (the engine): launches are recorded, not run.
  // Scans then use their own descriptors with device epochs (the call index
  // cap_epoch plus the base at cap_epoch_base, which the round's status
  // kernel advances); a buffer that would have to grow fails the capture
  // (the round then runs eagerly and grows it).
  bool capturing = false;
  unsigned cap_epoch = 0;
  const unsigned* cap_epoch_base = nullptr;
  Buf scan_desc_g;
  long long alloc_gen = 0;   // counts (re)allocations: a recorded graph names the buffers of its generation
  // run loop
  Buf RV[2], RF[2], keep, kpos, undo_in;
  // stage 1: distance keys, sample states, triangle votes; extraction ids,
  // per-cube offsets and the extracted mesh
  Buf rm_key, rm_st, rm_votes, rm_ids, rm_toff, rm_coff, rm_code, rm_tab, rm_V, rm_F;
  int rm_nV = 0, rm_nF = 0;
  int run_cur = 0, run_V = 0, run_F = 0;
  std::vector<int> run_faces;   // faces after each round of the last run
  int run_handle = -1;          // engine slot of the last pamo_run_stage2
  long long pair_cap_limit = -1;  // test hook: caps the traversal's pair buffer (-1: none)
  int guard_mode = 0;             // test hook: 0 by mesh size, 1 tree only, 2 exhaustive only
  int walk_split = -1;            // test hook: lanes per query of the LBVH walk, log2 (-1: by size)
  int strict = 0;                 // strict guard (engine; not in the reference)
  int placement = 0;              // 0 edge midpoint (the reference's), 1 optimal vertex (engine; not in the reference)
  Tuning tune;                    // launch shapes (results do not depend on them)

  int* hc = nullptr;   // pinned host mirror of the counters

  // No-allocation mode: every buffer is carved, 256-byte aligned, from one
  // fixed device region (arena) instead of hipMalloc; nothing is freed, and
  // the region is reused from its start by the next context set up on it.
  // The counter mirror is then the pageable hc_host.
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  int hc_host[C_COUNT] = {0};

  // optional per-kernel timing with HIP events (mask of PAMO_K_* bits)
  unsigned prof_mask = 0;
  struct ProfRec { int k; hipEvent_t a, b; double bytes; };
  std::vector<ProfRec> prof_open;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[32] = {0}, prof_bytes[32] = {0};
  long long prof_n[32] = {0};
  hipEvent_t take_event() {
    if (!ev_pool.empty()) { hipEvent_t e = ev_pool.back(); ev_pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  int prof_begin(int k, double bytes) {
    if (!(prof_mask & (1u << k))) return -1;
    ProfRec r{k, take_event(), take_event(), bytes};
    (void)hipEventRecord(r.a, st);
    prof_open.push_back(r);
    return (int)prof_open.size() - 1;
  }
  void prof_end(int h) {
    if (h < 0) return;
    (void)hipEventRecord(prof_open[h].b, st);
  }
  int prof_count() const { return (int)prof_open.size(); }
  // the last n records: dropped (their launches skipped themselves), or
  // given the byte counts of the sizes their round actually started from
  void prof_drop_last(int n) {
    for (; n > 0 && !prof_open.empty(); n--) {
      ev_pool.push_back(prof_open.back().a);
      ev_pool.push_back(prof_open.back().b);
      prof_open.pop_back();
    }
  }
  void prof_rebytes_last(int n, double (*bytes)(int, double, double), double F, double V) {
    for (int i = std::max(0, (int)prof_open.size() - n); i < (int)prof_open.size(); i++)
      prof_open[i].bytes = bytes(prof_open[i].k, F, V);
  }
  void prof_resolve() {
    for (auto& r : prof_open) {
      float ms = 0;
      (void)hipEventSynchronize(r.b);
      (void)hipEventElapsedTime(&ms, r.a, r.b);
      prof_ms[r.k] += ms;
      prof_bytes[r.k] += r.bytes;
      prof_n[r.k] += 1;
      ev_pool.push_back(r.a);
      ev_pool.push_back(r.b);
    }
    prof_open.clear();
  }

  // last forward
  int V = 0, F = 0, E = 0, n_fan = 0, n_vertices_undo = 0, n_winners = 0, n_collapsed = 0, cut = 0;
  int undo_passes = 0, n_checks = 0;
  uint32_t collapse_t_ = 0;
  long long n_queries = 0, n_cand = 0, n_hitfaces = 0;
  bool have_s_first = false;

  bool fail(const char* fmt, ...) {
    char b[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof(b), fmt, ap);
    va_end(ap);
    err = b;
    return false;
  }
  bool hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    return fail("%s: %s", what, hipGetErrorString(e));
  }
  bool check_launch(const char* what) { return hip(hipGetLastError(), what); }

  // nb bytes from the arena (no-allocation mode); the buffer's previous
  // slice, if any, stays unused until the region is set up again
  // every Buf member of the context (engine and stage-3 buffers are their own)
  std::vector<Buf*> bufs() {
    return {&pts_map, &table, &edge_cost, &undo_list, &tmp_undo, &P, &P0, &occ, &T, &T0, &cnt, &fan_off, &fan,
           &ecnt, &eoff, &edges, &Q, &tri_min, &win, &winpos, &collapsed, &qflag, &qf, &qpos, &af, &apos,
           &qlist, &pairs, &ovf, &seg_lo, &seg_hi, &leaf_face, &lbox_lo, &lbox_hi, &keys0, &keys1, &vals0, &vals1, &hist, &kleft, &kright,
           &kparent, &rfirst, &rlast, &nodeA, &nodeB, &S, &S_first, &mflag, &mpos, &marked,
           &scan_desc, &scan_desc_g, &counters, &cstats, &shuf_flags, &shuf_pos, &RV[0], &RV[1], &RF[0], &RF[1], &keep,
           &kpos, &undo_in, &rm_key, &rm_st, &rm_votes, &rm_ids, &rm_toff, &rm_coff, &rm_code, &rm_tab, &rm_V, &rm_F};
  }
  bool carve(Buf& b, size_t nb) {
    const size_t off = (arena_used + 255) & ~(size_t)255;
    if (off > arena_cap || nb > arena_cap - off)
      return fail("scratch region exhausted: %zu bytes in use, %zu more requested, %zu available", arena_used, nb,
                  arena_cap);
    b.p = arena + off;
    b.bytes = nb;
    arena_used = off + nb;
    return true;
  }
  // grow without preserving contents
  // Growth takes half again the request (an eighth in the fixed region of
  // the no-allocation entry point): a batch that fills from empty (serving)
  // regrows its round buffers a few times instead of dozens, and each
  // regrowth frees (which waits for the device) and maps memory.
  bool ensure(Buf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return true;
    if (capturing) return fail("buffer growth during graph capture");
    alloc_gen++;
    size_t nb = std::max<size_t>(bytes + bytes / 2, 256);
    if (arena) return carve(b, std::max<size_t>(bytes + bytes / 8, 256));   // the fixed region's own rule
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (!hip(hipMalloc(&b.p, nb), "hipMalloc")) return false;
    b.bytes = nb;
    return true;
  }
  // grow preserving contents (stream-ordered copy)
  bool ensure_keep(Buf& b, size_t bytes) {
    if (bytes <= b.bytes && b.p) return true;
    if (capturing) return fail("buffer growth during graph capture");
    alloc_gen++;
    Buf nbuf;
    size_t nb = std::max<size_t>(bytes + bytes / 2, 256);
    if (arena) {
      if (!carve(nbuf, nb)) return false;
    } else {
      if (!hip(hipMalloc(&nbuf.p, nb), "hipMalloc")) return false;
      nbuf.bytes = nb;
    }
    if (b.p) {
      if (!hip(hipMemcpyAsync(nbuf.p, b.p, b.bytes, hipMemcpyDeviceToDevice, st), "copy")) return false;
      if (!hip(hipStreamSynchronize(st), "sync")) return false;
      if (!arena) (void)hipFree(b.p);
    }
    b = nbuf;
    return true;
  }
  template <class T> T* ptr(Buf& b) { return (T*)b.p; }
  int* cnt_dev(int i) { return ptr<int>(counters) + i; }

  bool fill_i32(void* p, int v, long long n) {
    if (n <= 0) return true;
    k_fill_i32<<<nblk(n), BLOCK, 0, st>>>((int*)p, v, n);
    return check_launch("fill");
  }
  bool fill_u8(void* p, uint8_t v, long long n) {
    if (n <= 0) return true;
    k_fill_u8<<<nblk(n), BLOCK, 0, st>>>((uint8_t*)p, v, n);
    return check_launch("fill8");
  }
  bool set_counter(int i, int v) {
    k_fill_i32<<<1, 1, 0, st>>>(cnt_dev(i), v, 1);
    return check_launch("set counter");
  }
  // exclusive scan of n ints (non-negative); in-place allowed; total ->
  // counter slot (or -1).  One launch (decoupled look-back).
  bool scan(const int* in, int* out, long long n, int total_slot) {
    if (n <= 0) {
      if (total_slot >= 0) return set_counter(total_slot, 0);
      return true;
    }
    return scan_launch(in, out, n, nullptr, 0, total_slot >= 0 ? cnt_dev(total_slot) : nullptr);
  }
  // scan of 0/1 flags that also writes the index of every set flag at its
  // position in list (stream compaction in the same pass)
  bool scan_compact(const int* in, int* out, long long n, int total_slot, int* list) {
    if (n <= 0) {
      if (total_slot >= 0) return set_counter(total_slot, 0);
      return true;
    }
    return scan_launch(in, out, n, nullptr, 0, total_slot >= 0 ? cnt_dev(total_slot) : nullptr, list);
  }
  // the same scan over a capacity n whose live length is on the device
  // (*d_n + add); entries past the live length are neither read nor written
  bool scan_dev(const int* in, int* out, long long n, const int* d_n, int add, int* total = nullptr,
                int* list = nullptr, const int* gate = nullptr) {
    if (n <= 0) return true;
    return scan_launch(in, out, n, d_n, add, total, list, gate);
  }
  bool scan_launch(const int* in, int* out, long long n, const int* d_n, int add, int* total, int* list = nullptr,
                   const int* gate = nullptr) {
    return scan_op(ScanArray{in, out, list}, n, d_n, add, total, SCAN_SMALL_MAX, gate);
  }
  // one-pass scan whose input and output are the op's count and emit
  // (inputs up to small_max take the small tile shape; a non-zero *gate on
  // the device skips the scan)
  template <class Op>
  bool scan_op(const Op& op, long long n, const int* d_n, int add, int* total, long long small_max = SCAN_SMALL_MAX,
               const int* gate = nullptr) {
    if (n <= 0) return true;
    const bool small = n <= small_max;
    const long long tile = small ? SCANS_TILE : SCAN1_TILE;
    long long nb = (n + tile - 1) / tile;
    if (capturing) {
      // device epochs over the graph's own descriptors (sized and cleared
      // before the capture)
      if (scan_desc_g.bytes < (size_t)nb * 8 || cap_epoch + 1 >= SCAN_EPOCH_STEP)
        return fail("graph capture: scan descriptors");
      const unsigned ep = ++cap_epoch;
      unsigned long long* dg = ptr<unsigned long long>(scan_desc_g);
      if (small)
        k_scan_op<Op, SCANS_NT, SCANS_ITEMS><<<(unsigned)nb, SCANS_NT, 0, st>>>(
            op, n, d_n, add, dg, ep, cap_epoch_base, total, cnt_dev(C_ERR), gate);
      else
        k_scan_op<Op><<<(unsigned)nb, BLOCK1K, 0, st>>>(op, n, d_n, add, dg, ep, cap_epoch_base, total,
                                                        cnt_dev(C_ERR), gate);
      return check_launch("scan");
    }
    // descriptors start at zero (unpublished); cleared by a fill kernel
    if (scan_desc.bytes < (size_t)nb * 8) {
      if (!ensure(scan_desc, (size_t)nb * 8)) return false;
      if (!fill_i32(scan_desc.p, 0, (long long)(scan_desc.bytes / 4))) return false;
    }
    if (++scan_epoch >= (1u << 22)) {   // epoch wrap: clear the descriptors
      scan_epoch = 1;
      if (!fill_i32(scan_desc.p, 0, (long long)(scan_desc.bytes / 4))) return false;
    }
    if (small)
      k_scan_op<Op, SCANS_NT, SCANS_ITEMS><<<(unsigned)nb, SCANS_NT, 0, st>>>(
          op, n, d_n, add, ptr<unsigned long long>(scan_desc), scan_epoch, nullptr, total, cnt_dev(C_ERR), gate);
    else
      k_scan_op<Op><<<(unsigned)nb, BLOCK1K, 0, st>>>(op, n, d_n, add, ptr<unsigned long long>(scan_desc), scan_epoch,
                                                      nullptr, total, cnt_dev(C_ERR), gate);
    return check_launch("scan");
  }
  bool read_counters() {
    if (!hip(hipMemcpyAsync(hc, counters.p, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st), "read counters"))
      return false;
    return hip(hipStreamSynchronize(st), "sync");
  }

  bool ensure_pts_storage(size_t n) {   // cusimp_free.cu:47-85 growth rule
    if (n <= alloc_pts) return true;
    alloc_pts = n + n / 5;
    if (!ensure(pts_map, (alloc_pts + 1) * 4) || !ensure(table, alloc_pts + 1)) return false;
    if (!fill_i32(pts_map.p, 0, (long long)alloc_pts + 1) || !fill_u8(table.p, 0, (long long)alloc_pts + 1))
      return false;
    // the undo lists are reallocated too: their contents are lost (D1)
    if (!ensure_keep(undo_list, (alloc_pts + 1) * 4) || !ensure_keep(tmp_undo, (alloc_pts + 1) * 4)) return false;
    if (!fill_i32(undo_list.p, 0, (long long)(undo_list.bytes / 4)) || !fill_i32(tmp_undo.p, 0, (long long)(tmp_undo.bytes / 4)))
      return false;
    return true;
  }
  bool ensure_edge_cost_storage(size_t n) {   // cusimp_free.cu:180-194
    if (n <= alloc_edge_cost) return true;
    alloc_edge_cost = n + n / 5;
    if (!ensure(edge_cost, (alloc_edge_cost + 1) * 4)) return false;
    return fill_i32(edge_cost.p, 0, (long long)alloc_edge_cost + 1);
  }

  bool build_bvh(int Fn);
  bool build_bvh_and_traverse(int Fn, bool debug_pairs, std::vector<int>* pairs_out);
  bool node_boxes(const int* n_alive, const int* sorted_leaf, long long Fl);
  bool check_and_mark(int* n_marked);
  int forward(const float* d_pts, const int* d_tris, const int* d_undo, int n_u, int nV, int nF, float scale,
              float thr, int is_stuck, int init, int sp = 0);
  ~pamo_ctx();
};

static void eng_free(pamo_ctx* c);
static void s3_free(pamo_ctx* c);

pamo_ctx::~pamo_ctx() {
  (void)hipSetDevice(device);
  if (arena) {   // no-allocation mode: the region is not ours to free
    eng_free(this);
    s3_free(this);
    return;
  }
  (void)hipDeviceSynchronize();
  eng_free(this);
  s3_free(this);
  for (Buf* b : bufs())
    if (b->p) (void)hipFree(b->p);
  for (auto& r : prof_open) { ev_pool.push_back(r.a); ev_pool.push_back(r.b); }
  for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  if (hc) (void)hipHostFree(hc);
}

// Node boxes of the LBVH over n_alive sorted leaves (capacity Fl): in-block
// merge plus segment levels 1..8, higher segment levels, then the nodes that
// span blocks.
bool pamo_ctx::node_boxes(const int* n_alive, const int* sorted_leaf, long long Fl) {
  if (!ensure(seg_lo, (size_t)seg_entries(Fl) * 16) || !ensure(seg_hi, (size_t)seg_entries(Fl) * 16)) return false;
  k_bottom_up_local<<<nblk(Fl), BLOCK, 0, st>>>(n_alive, ptr<float4>(lbox_lo), ptr<float4>(lbox_lo) + 1, sorted_leaf,
                                               ptr<int>(kparent), ptr<int>(kleft),
                                               ptr<int>(kright), ptr<int>(rfirst), ptr<int>(rlast), ptr<float4>(nodeA),
                                               ptr<float4>(nodeB), Fl, ptr<float4>(seg_lo), ptr<float4>(seg_hi));
  for (int j0 = 8; (1LL << j0) < Fl; j0 += 8) {
    long long entries = (Fl + (1LL << j0) - 1) >> j0;
    k_seg_upper<<<nblk(entries), BLOCK, 0, st>>>(n_alive, Fl, j0, ptr<float4>(seg_lo), ptr<float4>(seg_hi));
  }
  k_cross_nodes<<<nblk(Fl), BLOCK, 0, st>>>(n_alive, Fl, ptr<int>(kleft), ptr<int>(rfirst), ptr<int>(rlast),
                                           ptr<float4>(seg_lo), ptr<float4>(seg_hi), ptr<float4>(nodeA),
                                           ptr<float4>(nodeB));
  return check_launch("node boxes");
}

// LBVH over the faces leaf_face[0..n) (n at C_NALIVE, at most Fn) of P/T:
// nodes in nodeA/nodeB.
bool pamo_ctx::build_bvh(int Fn) {
  const long long Fl = std::max(Fn, 1);
  if (!ensure(lbox_lo, Fl * 32) || !ensure(lbox_hi, Fl * 16) || !ensure(keys0, Fl * 4) || !ensure(keys1, Fl * 4) ||
      !ensure(vals0, Fl * 4) || !ensure(vals1, Fl * 4) || !ensure(kleft, Fl * 4) || !ensure(kright, Fl * 4) ||
      !ensure(kparent, 2 * Fl * 4) || !ensure(rfirst, Fl * 4) || !ensure(rlast, Fl * 4) ||
      !ensure(nodeA, 2 * Fl * 16) || !ensure(nodeB, 2 * Fl * 16))
    return false;
  int* n_alive = cnt_dev(C_NALIVE);
  int* scene = ptr<int>(cstats) + 8;
  // LBVH build bytes: alive faces counted as F (the alive count is on the device)
  int pbv = prof_begin(PAMO_K_LBVH, 12.0 * Fn + 12.0 * V + 8.0 * Fn + 8.0 * Fn + 8.0 * Fn + (2.0 * Fn - 1) * 16 +
                                       (2.0 * Fn - 1) * 24);
  k_scene_init<<<1, 64, 0, st>>>(scene);
  k_leaf_boxes<<<nblk(Fl), BLOCK, 0, st>>>(ptr<int>(leaf_face), n_alive, ptr<float>(P), ptr<int>(T),
                                          ptr<float4>(lbox_lo), ptr<float4>(lbox_lo) + 1, scene);
  k_morton<<<nblk(Fl), BLOCK, 0, st>>>(ptr<float4>(lbox_lo), ptr<float4>(lbox_lo) + 1, n_alive, scene,
                                      ptr<uint32_t>(keys0), ptr<int>(vals0));
  if (!check_launch("leaf boxes")) return false;
  // radix sort: 4 passes of 8 bits over 30-bit Morton codes
  long long ntiles = (Fl + RS_TILE - 1) / RS_TILE;
  if (!ensure(hist, (size_t)ntiles * 256 * 4)) return false;
  uint32_t* kin = ptr<uint32_t>(keys0);
  uint32_t* kout = ptr<uint32_t>(keys1);
  int* vin = ptr<int>(vals0);
  int* vout = ptr<int>(vals1);
  for (int shift = 0; shift < 32; shift += 8) {
    k_rs_hist<<<(unsigned)ntiles, BLOCK, 0, st>>>(kin, n_alive, shift, (int)ntiles, ptr<int>(hist));
    if (!scan(ptr<int>(hist), ptr<int>(hist), ntiles * 256, -1)) return false;
    k_rs_scatter<<<(unsigned)ntiles, BLOCK, 0, st>>>(kin, vin, n_alive, shift, (int)ntiles, ptr<int>(hist), kout,
                                                     vout);
    if (!check_launch("radix sort")) return false;
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  k_fill_i32<<<1, 1, 0, st>>>(ptr<int>(kparent), -1, 1);
  k_karras<<<nblk(Fl), BLOCK, 0, st>>>(kin, n_alive, ptr<int>(kleft), ptr<int>(kright), ptr<int>(kparent),
                                      ptr<int>(rfirst), ptr<int>(rlast));
  if (!node_boxes(n_alive, vin, Fl)) return false;
  prof_end(pbv);
  return check_launch("lbvh");
}

// BVH over the alive faces, then the restricted traversal (query list in
// qlist/C_NQUERY) or, for the debug export, the full pair list.
bool pamo_ctx::build_bvh_and_traverse(int Fn, bool debug_pairs, std::vector<int>* pairs_out) {
  const long long Fl = std::max(Fn, 1);
  if (!build_bvh(Fn)) return false;
  int* n_alive = cnt_dev(C_NALIVE);
  if (!debug_pairs) {
    int pt = prof_begin(PAMO_K_TRAVERSE, 0.0);
    if (!fill_u8(S.p, 0, Fl)) return false;
    k_traverse<<<nblk(Fl), BLOCK, 0, st>>>(ptr<int>(qlist), cnt_dev(C_NQUERY), n_alive, ptr<float4>(nodeA),
                                          ptr<float4>(nodeB), ptr<float>(P), ptr<int>(T), ptr<uint8_t>(S),
                                          (unsigned long long*)cstats.p, cnt_dev(C_ERR));
    prof_end(pt);
    return check_launch("traverse");
  }
  // debug: count, scan, emit
  if (!ensure(qpos, (Fl + 1) * 4) || !ensure(mflag, (Fl + 1) * 4)) return false;
  if (!fill_i32(mflag.p, 0, Fl + 1)) return false;
  k_traverse_list<<<nblk(Fl), BLOCK, 0, st>>>(ptr<int>(qlist), cnt_dev(C_NQUERY), ptr<float4>(nodeA),
                                             ptr<float4>(nodeB), ptr<float>(P), ptr<int>(T), ptr<int>(mflag), nullptr,
                                             nullptr);
  if (!scan(ptr<int>(mflag), ptr<int>(qpos), Fl + 1, C_NMARK)) return false;
  if (!read_counters()) return false;
  long long np = hc[C_NMARK];
  Buf pb;
  if (!hip(hipMalloc(&pb.p, std::max<long long>(np, 1) * 12), "hipMalloc pairs")) return false;
  k_traverse_list<<<nblk(Fl), BLOCK, 0, st>>>(ptr<int>(qlist), cnt_dev(C_NQUERY), ptr<float4>(nodeA),
                                             ptr<float4>(nodeB), ptr<float>(P), ptr<int>(T), nullptr, ptr<int>(qpos),
                                             (int*)pb.p);
  pairs_out->resize(3 * np);
  bool ok = check_launch("pairs") &&
            hip(hipMemcpyAsync(pairs_out->data(), pb.p, np * 12, hipMemcpyDeviceToHost, st), "copy pairs") &&
            hip(hipStreamSynchronize(st), "sync");
  (void)hipFree(pb.p);
  return ok;
}

// One self-intersection check restricted to the fans of the first `cut`
// collapsed edges, then the undo marks (cusimp_free.cu:1014-1022).
bool pamo_ctx::check_and_mark(int* n_marked) {
  n_checks++;
  const long long Fl = std::max(F, 1);
  if (!ensure(qflag, Fl) || !ensure(qf, Fl * 4) || !ensure(qpos, (Fl + 1) * 4) || !ensure(af, Fl * 4) ||
      !ensure(apos, Fl * 4) || !ensure(qlist, Fl * 4) || !ensure(leaf_face, Fl * 4) || !ensure(S, Fl))
    return false;
  if (!fill_u8(qflag.p, 0, Fl)) return false;
  k_qflag_mark<<<nblk(cut), BLOCK, 0, st>>>(ptr<int>(collapsed), cnt_dev(C_CUT), ptr<int>(edges), ptr<int>(fan_off),
                                           ptr<int>(fan), ptr<uint8_t>(qflag));
  k_query_flags<<<nblk(F), BLOCK, 0, st>>>(ptr<uint8_t>(qflag), ptr<int>(T), F, ptr<int>(qf), ptr<int>(af));
  if (!check_launch("query flags")) return false;
  if (!scan(ptr<int>(qf), ptr<int>(qpos), F, C_NQUERY)) return false;
  k_compact_idx<<<nblk(F), BLOCK, 0, st>>>(ptr<int>(qf), ptr<int>(qpos), F, ptr<int>(qlist), 0);
  if (!scan(ptr<int>(af), ptr<int>(apos), F, C_NALIVE)) return false;
  k_compact_idx<<<nblk(F), BLOCK, 0, st>>>(ptr<int>(af), ptr<int>(apos), F, ptr<int>(leaf_face), 0);
  if (!check_launch("compact")) return false;
  if (!build_bvh_and_traverse(F, false, nullptr)) return false;
  // marks
  if (!ensure(mflag, ((size_t)cut + 1) * 4) || !ensure(mpos, ((size_t)cut + 1) * 4) ||
      !ensure(marked, ((size_t)cut + 1) * 4) || !ensure_keep(tmp_undo, ((size_t)2 * cut + 2) * 4))
    return false;
  k_mark_undo<<<nblk(cut), BLOCK, 0, st>>>(ptr<int>(collapsed), cnt_dev(C_CUT), ptr<int>(edges), ptr<int>(fan_off),
                                          ptr<int>(fan), ptr<uint8_t>(S), ptr<int>(mflag));
  if (!scan(ptr<int>(mflag), ptr<int>(mpos), cut, C_NMARK)) return false;
  k_marked_pairs<<<nblk(cut), BLOCK, 0, st>>>(ptr<int>(mflag), ptr<int>(mpos), cnt_dev(C_CUT), ptr<int>(collapsed),
                                             ptr<int>(edges), ptr<int>(marked), ptr<int>(tmp_undo));
  if (!check_launch("marks")) return false;
  if (!read_counters()) return false;
  if (hc[C_ERR]) return fail("malformed LBVH (%d walks exceeded the node count)", hc[C_ERR]);
  *n_marked = hc[C_NMARK];
  n_queries += hc[C_NQUERY];
  return true;
}

int pamo_ctx::forward(const float* d_pts, const int* d_tris, const int* d_undo, int n_u, int nV, int nF, float sc,
                      float thr, int is_stuck, int init, int sp) {
  err.clear();
  if (sp && (n_u != 0 || is_stuck)) { fail("the unguarded round takes no undo list"); return -1; }
  if (nV < 0 || nF < 0 || n_u < 0) { fail("negative size"); return -1; }
  if (!hip(hipSetDevice(device), "hipSetDevice")) return -1;
  if (!ensure(counters, C_COUNT * 4) || !ensure(cstats, 64 * 4)) return -1;
  if (!fill_i32(counters.p, 0, C_COUNT) || !fill_i32(cstats.p, 0, 8)) return -1;
  n_checks = 0; undo_passes = 0; n_queries = 0; have_s_first = false;

  // step 1 (cusimp_free.cu:895-916) on the storage as it stands
  n_vertices_undo = 0;
  int first = 0;
  if (is_stuck) {
    if (n_u > 0) {
      if (!ensure_keep(tmp_undo, ((size_t)n_u + 1) * 4) || !ensure_keep(undo_list, ((size_t)n_u + 1) * 4)) return -1;
      if (!hip(hipMemcpyAsync(tmp_undo.p, d_undo, (size_t)n_u * 4, hipMemcpyDeviceToDevice, st), "copy undo"))
        return -1;
      long long map_n = alloc_pts ? (long long)alloc_pts + 1 : 0;
      k_remap_undo<<<nblk(n_u), BLOCK, 0, st>>>(ptr<int>(tmp_undo), ptr<int>(pts_map), map_n, ptr<int>(undo_list), 0,
                                               n_u, cnt_dev(C_ERR));
      if (!check_launch("remap undo")) return -1;
    }
    n_vertices_undo += n_u;
    first += n_u;
  } else if (n_u != 0 && alloc_pts) {
    if (!fill_u8(table.p, 0, (long long)alloc_pts + 1)) return -1;
  }
  if (n_u != 0) {
    long long tn = alloc_pts ? (long long)alloc_pts + 1 : 0;
    k_set_table<<<nblk(n_u), BLOCK, 0, st>>>(d_undo, n_u, ptr<uint8_t>(table), tn, cnt_dev(C_ERR));
    if (!check_launch("table")) return -1;
  }
  const float tau = thr;
  collapse_t_ = collapse_threshold(thr);
  V = nV;
  F = nF;
  if (!ensure_pts_storage(V)) return -1;
  const long long Vl = std::max(V, 1), Fl = std::max(F, 1);
  if (!ensure(P, Vl * 12) || !ensure(P0, Vl * 12) || !ensure(occ, (Vl + 1) * 4) || !ensure(T, Fl * 12) ||
      !ensure(T0, Fl * 12))
    return -1;
  k_validate<<<nblk(F), BLOCK, 0, st>>>(d_tris, F, V, cnt_dev(C_ERR));
  k_init_points<<<nblk(V + 1), BLOCK, 0, st>>>(d_pts, ptr<float>(P), ptr<float>(P0), ptr<int>(occ), V);
  if (!check_launch("init points")) return -1;
  if (init && shuffle_mode == PAMO_SHUFFLE_FEISTEL && F > 0) {
    Feistel fb;
    uint64_t m = (uint64_t)F, total;
    if (m <= 16) total = 4;
    else { total = 0; uint64_t t = m - 1; while (t) { total++; t >>= 1; } }
    fb.lbits = (uint32_t)(total / 2);
    fb.lmask = (uint32_t)((1ull << fb.lbits) - 1);
    fb.rbits = (uint32_t)(total - fb.lbits);
    fb.rmask = (uint32_t)((1ull << fb.rbits) - 1);
    uint64_t x = 1;
    for (int i = 0; i < 24; i++) { x = (x * 48271ull) % 2147483647ull; fb.key[i] = (uint32_t)x; }
    long long n = 1LL << (fb.lbits + fb.rbits);
    if (!ensure(shuf_flags, n * 4) || !ensure(shuf_pos, n * 4)) return -1;
    k_shuffle_flags<<<nblk(n), BLOCK, 0, st>>>(fb, n, F, ptr<int>(shuf_flags));
    if (!scan(ptr<int>(shuf_flags), ptr<int>(shuf_pos), n, -1)) return -1;
    k_shuffle_scatter<<<nblk(n), BLOCK, 0, st>>>(fb, n, F, ptr<int>(shuf_pos), d_tris, ptr<int>(T), ptr<int>(T0));
  } else {
    k_init_tris<<<nblk(3LL * F), BLOCK, 0, st>>>(d_tris, ptr<int>(T), ptr<int>(T0), F);
  }
  if (!check_launch("init tris")) return -1;

  // fans and edges (cusimp_free.cu:950-978)
  if (!ensure(cnt, (Vl + 1) * 4) || !ensure(fan_off, (Vl + 1) * 4) || !ensure(ecnt, (Fl + 1) * 4) ||
      !ensure(eoff, (Fl + 1) * 4))
    return -1;
  if (!fill_i32(cnt.p, 0, V + 1)) return -1;
  k_fan_count<<<nblk(F), BLOCK, 0, st>>>(ptr<int>(T), F, ptr<int>(cnt));
  if (!scan(ptr<int>(cnt), ptr<int>(fan_off), V + 1, C_NFAN)) return -1;
  k_edge_count<<<nblk(F + 1), BLOCK, 0, st>>>(ptr<int>(T), F, ptr<int>(ecnt));
  if (!scan(ptr<int>(ecnt), ptr<int>(eoff), F + 1, C_E)) return -1;
  if (!read_counters()) return -1;
  if (hc[C_ERR]) { fail("invalid input: %d face or undo indices out of range", hc[C_ERR]); return -2; }
  E = hc[C_E];
  n_fan = hc[C_NFAN];
  const long long El = std::max(E, 1);
  if (!ensure(fan, (size_t)std::max(n_fan, 1) * 4) || !ensure(edges, El * 8) || !ensure(Q, Vl * 40) ||
      !ensure(tri_min, Fl * 8) || !ensure(win, (El + 1) * 4) || !ensure(winpos, (El + 1) * 4) ||
      !ensure(collapsed, (2 * El + 1) * 4))
    return -1;
  if (!fill_i32(cnt.p, 0, V + 1)) return -1;
  k_fan_fill<<<nblk(F), BLOCK, 0, st>>>(ptr<int>(T), F, ptr<int>(fan_off), ptr<int>(cnt), ptr<int>(fan));
  k_fan_sort<<<nblk(V), BLOCK, 0, st>>>(ptr<int>(fan_off), ptr<int>(fan), V);
  k_edge_emit<<<nblk(F), BLOCK, 0, st>>>(ptr<int>(T), F, ptr<int>(eoff), ptr<int>(edges));
  // quadrics, costs, claims, collapse (cusimp_free.cu:980-999)
  int ph = prof_begin(PAMO_K_QUADRIC, 12.0 * F + 12.0 * V + 4.0 * (V + 1) + 12.0 * F + 40.0 * V);
  k_quadric<<<nblk(V), BLOCK, 0, st>>>(ptr<float>(P), ptr<int>(T), ptr<int>(fan_off), ptr<int>(fan), V, ptr<float>(Q));
  prof_end(ph);
  if (!check_launch("fans/edges/quadrics")) return -1;
  if (!ensure_edge_cost_storage(E)) return -1;
  ph = prof_begin(PAMO_K_EDGE_COST, 8.0 * E + 12.0 * V + 40.0 * V + 4.0 * (V + 1) + 12.0 * F + 12.0 * F + 1.0 * V +
                                        4.0 * E);
  k_edge_cost<<<nblk(E), BLOCK, 0, st>>>(ptr<float>(P), ptr<int>(T), ptr<int>(fan_off), ptr<int>(fan), ptr<int>(edges),
                                        E, ptr<float>(Q), ptr<uint8_t>(table), ptr<uint32_t>(edge_cost), sc, tau,
                                        is_stuck, sp);
  prof_end(ph);
  k_fill_u64<<<nblk(F), BLOCK, 0, st>>>(ptr<u64>(tri_min), ~0ull, F);
  ph = prof_begin(PAMO_K_CLAIM, 4.0 * E + 8.0 * E + 4.0 * (V + 1) + 12.0 * F + 8.0 * F);
  k_claim<<<nblk(E), BLOCK, 0, st>>>(ptr<int>(edges), E, ptr<uint32_t>(edge_cost), ptr<int>(fan_off), ptr<int>(fan),
                                    ptr<u64>(tri_min));
  prof_end(ph);
  k_collapse<<<nblk(E), BLOCK, 0, st>>>(ptr<int>(edges), E, ptr<uint32_t>(edge_cost), collapse_t_, ptr<int>(fan_off),
                                       ptr<int>(fan), ptr<u64>(tri_min), ptr<float>(P), ptr<float>(P0), ptr<int>(T),
                                       ptr<int>(T0), ptr<int>(occ), ptr<int>(win), sp);
  if (!check_launch("cost/claim/collapse")) return -1;
  if (sp) {   // the unguarded round ends here (cusimp.cu:713-716): the vertex map
    n_winners = n_collapsed = cut = 0;
    undo_passes = 0;
    if (!scan(ptr<int>(occ), ptr<int>(pts_map), V + 1, C_NEWV) || !read_counters()) return -1;
    n_cand = n_hitfaces = 0;
    return 0;
  }
  if (!scan(ptr<int>(win), ptr<int>(winpos), E, C_NWIN)) return -1;
  k_compact_idx<<<nblk(E), BLOCK, 0, st>>>(ptr<int>(win), ptr<int>(winpos), E, ptr<int>(collapsed), 0);
  k_remove_invalid_faces<<<nblk(F), BLOCK, 0, st>>>(ptr<int>(T), F);
  k_line_detect<<<nblk(E), BLOCK, 0, st>>>(ptr<float>(P), ptr<int>(edges), E, cnt_dev(C_NLINE));
  if (!check_launch("collapse list")) return -1;
  if (!read_counters()) return -1;
  n_winners = hc[C_NWIN];
  n_collapsed = n_winners;
  if (hc[C_NLINE] > 0) {
    // the sequential pass may append up to E entries
    if (!set_counter(C_NCOLL, n_winners)) return -1;
    k_line_apply_seq<<<1, 1, 0, st>>>(ptr<float>(P), ptr<int>(T), ptr<int>(occ), ptr<int>(edges), E, ptr<int>(fan_off),
                                     ptr<int>(fan), ptr<int>(collapsed), cnt_dev(C_NCOLL));
    if (!check_launch("line edges") || !read_counters()) return -1;
    n_collapsed = hc[C_NCOLL];
  }
  cut = n_winners == 0 ? 0 : std::min(n_collapsed, ((n_winners + 511) / 512) * 512);
  if (!set_counter(C_CUT, cut)) return -1;

  // self-intersection guard and undo loop (cusimp_free.cu:1014-1040)
  int hm = 0;
  if (cut > 0) {
    if (!check_and_mark(&hm)) return -1;
    if (!ensure(S_first, Fl) ||
        !hip(hipMemcpyAsync(S_first.p, S.p, F, hipMemcpyDeviceToDevice, st), "copy S"))
      return -1;
    have_s_first = true;
  }
  n_vertices_undo += 2 * hm;
  int it = 0;
  while (hm != 0) {
    it++;
    k_fill_i32<<<1, 1, 0, st>>>(cnt_dev(C_NMARK), hm, 1);
    k_undo<<<nblk(hm), BLOCK, 0, st>>>(ptr<int>(marked), cnt_dev(C_NMARK), ptr<int>(edges), ptr<float>(P),
                                      ptr<float>(P0), ptr<int>(T), ptr<int>(T0), ptr<int>(occ), ptr<int>(fan_off),
                                      ptr<int>(fan));
    if (!ensure_keep(undo_list, ((size_t)n_vertices_undo + 1) * 4)) return -1;
    long long map_n = (long long)alloc_pts + 1;
    k_remap_undo<<<nblk(n_vertices_undo - first), BLOCK, 0, st>>>(ptr<int>(tmp_undo), ptr<int>(pts_map), map_n,
                                                                  ptr<int>(undo_list), first, n_vertices_undo - first,
                                                                  cnt_dev(C_ERR));
    if (!check_launch("undo")) return -1;
    first += 2 * hm;
    if (!check_and_mark(&hm)) return -1;
    n_vertices_undo += 2 * hm;
    if (it == 5) break;
  }
  undo_passes = it;
  if (!ensure_keep(undo_list, ((size_t)n_vertices_undo + 1) * 4)) return -1;
  if (n_vertices_undo > first && !fill_i32(ptr<int>(undo_list) + first, 0, n_vertices_undo - first)) return -1;   // D6
  // pts_map = exclusive scan of occupancy (cusimp_free.cu:1042-1045)
  if (!scan(ptr<int>(occ), ptr<int>(pts_map), V + 1, C_NEWV)) return -1;
  unsigned long long cs[2] = {0, 0};
  if (!hip(hipMemcpyAsync(cs, cstats.p, 16, hipMemcpyDeviceToHost, st), "stats") || !hip(hipStreamSynchronize(st), "sync"))
    return -1;
  n_cand = (long long)cs[0];
  n_hitfaces = (long long)cs[1];
  return 0;
}

#include "pamo_engine.inc"
#include "pamo_remesh.inc"
#include "pamo_stage3.inc"

// every Buf member of Engine
static std::vector<Buf*> eng_bufs(Engine& g) {
  return {&g.P, &g.P0, &g.T, &g.T0, &g.vmesh, &g.fmesh, &g.nP, &g.nT, &g.nvm, &g.nfm, &g.undo_in, &g.undo_in2,
         &g.undo_out, &g.tmp, &g.marked, &g.pm, &g.table, &g.ecost, &g.pm2, &g.table2, &g.ecost2, &g.roff,
         &g.mp, &g.voff, &g.foff, &g.eoffm, &g.uioff, &g.uioff2, &g.nustk, &g.h, &g.lcnt, &g.ncoll, &g.cut,
         &g.cutoff, &g.uobase, &g.hm, &g.part, &g.first, &g.nvu, &g.nvoff, &g.nfoff, &g.seg_a, &g.seg_b,
         &g.root, &g.root_end, &g.scene, &g.g, &g.keep, &g.kpos, &g.newk, &g.vkeep, &g.vpos, &g.npass,
         &g.fsucc, &g.sface, &g.splane, &g.slen, &g.erank, &g.vedge, &g.uoff, &g.vmin, &g.wlist, &g.frank,
         &g.ucnt, &g.oflag, &g.bfnb, &g.bfoff, &g.pk, &g.sorder, &g.sorder2, &g.nsort, &g.tbox, &g.bfq,
         &g.bfw, &g.bfl, &g.bfn, &g.ms, &g.live, &g.pkd[0], &g.pkd[1], &g.dbox, &g.kslot};
}

static void eng_free(pamo_ctx* c) {
  if (!c->eng) return;
  Engine& g = *c->eng;

  if (!c->arena) {
    for (Buf* b : eng_bufs(g))
      if (b->p) (void)hipFree(b->p);
    for (auto& s : g.slots) {
      if (s.outV.p) (void)hipFree(s.outV.p);
      if (s.outF.p) (void)hipFree(s.outF.p);
    }
    if (g.hb) (void)hipHostFree(g.hb);
    for (int i = 0; i < 2; i++) {
      if (g.hd[i]) (void)hipHostFree(g.hd[i]);
      if (g.dev_done[i]) (void)hipEventDestroy(g.dev_done[i]);
    }
    for (auto& e : g.graphs) (void)hipGraphExecDestroy(e.exec);
    if (g.cap_st) (void)hipStreamDestroy(g.cap_st);
  }
  delete c->eng;
  c->eng = nullptr;
}

static bool eng_ready(pamo_ctx& c) {
  if (!c.hip(hipSetDevice(c.device), "hipSetDevice")) return false;
  if (!c.eng) c.eng = new Engine();
  return c.ensure(c.counters, C_COUNT * 4) && c.ensure(c.cstats, 64 * 4);
}

// ====================================================================== ABI
extern "C" {

int pamo_batch_submit(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts, int n_faces,
                      float scale, long long target, float threshold, int tolerance, long long iters, long long job) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  c.err.clear();
  if (n_verts < 0 || n_faces < 0 || tolerance < 0 || iters < 1) { c.fail("invalid arguments"); return -1; }
  if (!eng_ready(c)) return -1;
  // faces must index [0, n_verts) (rows of -1 are removed faces)
  if (!c.fill_i32(c.counters.p, 0, C_COUNT)) return -1;
  k_validate<<<nblk(n_faces), BLOCK, 0, c.st>>>(d_faces, n_faces, n_verts, c.cnt_dev(C_ERR));
  if (!c.check_launch("validate") || !c.read_counters()) return -1;
  if (c.hc[C_ERR]) { c.fail("invalid input: %d face indices out of range", c.hc[C_ERR]); return -2; }
  // stream-ordered from here: the copies of the caller's buffers are
  // enqueued on the stream and complete before any later work on it
  return eng_admit(c, d_verts, d_faces, n_verts, n_faces, scale, target, threshold, tolerance, iters, job);
}

int pamo_batch_step(pamo_ctx* ctx, void* stream, int max_rounds, int* n_active) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  c.err.clear();
  if (!eng_ready(c)) return -1;
  int finished = 0;
  if (!eng_rounds(c, max_rounds, -1, &finished)) return -1;
  if (n_active) *n_active = c.eng->M;
  return finished;
}

int pamo_batch_poll(pamo_ctx* ctx, int* handles, long long* jobs, int cap) {
  if (!ctx || !ctx->eng) return 0;
  int n = 0;
  for (int h = 0; h < (int)ctx->eng->slots.size(); h++) {
    const EngSlot& s = ctx->eng->slots[h];
    if (!s.done || s.taken) continue;
    if (n < cap) {
      if (handles) handles[n] = h;
      if (jobs) jobs[n] = s.job;
    }
    n++;
  }
  return n;
}

int pamo_batch_result(pamo_ctx* ctx, int handle, int* n_verts, int* n_faces, long long* stats, int* round_faces,
                      int cap) {
  if (!ctx || !ctx->eng || handle < 0 || handle >= (int)ctx->eng->slots.size()) return -1;
  const EngSlot& s = ctx->eng->slots[handle];
  if (!s.done) return 1;
  if (n_verts) *n_verts = s.rV;
  if (n_faces) *n_faces = s.rF;
  if (stats) {
    long long v[8] = {s.rounds, s.exit_reason, s.collapses, s.max_passes, (long long)s.job, s.V0, s.F0,
                      (long long)((s.t_done - s.t_admit) * 1e6)};
    memcpy(stats, v, sizeof(v));
  }
  if (round_faces)
    for (int i = 0; i < (int)s.round_faces.size() && i < cap; i++) round_faces[i] = s.round_faces[i];
  return 0;
}

int pamo_batch_take(pamo_ctx* ctx, void* stream, int handle, float* d_verts_out, int* d_faces_out) {
  if (!ctx || !ctx->eng || handle < 0 || handle >= (int)ctx->eng->slots.size()) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  EngSlot& s = c.eng->slots[handle];
  if (!s.done || s.taken) { c.fail("job %d is not finished or already taken", handle); return -1; }
  if (s.rV && d_verts_out &&
      !c.hip(hipMemcpyAsync(d_verts_out, s.outV.p, (size_t)s.rV * 12, hipMemcpyDeviceToDevice, c.st), "take"))
    return -1;
  if (s.rF && d_faces_out &&
      !c.hip(hipMemcpyAsync(d_faces_out, s.outF.p, (size_t)s.rF * 12, hipMemcpyDeviceToDevice, c.st), "take"))
    return -1;
  if (!c.hip(hipStreamSynchronize(c.st), "sync")) return -1;
  if (!c.arena) {
    (void)hipFree(s.outV.p);
    (void)hipFree(s.outF.p);
  }
  s.outV = Buf();
  s.outF = Buf();
  s.taken = true;
  return 0;
}

// The pamo_blender entry point (pamo_blender/cusimpinterface.cu:57-161):
// host arrays in and out, one process-wide context, the PaMO.run loop with
// tolerance 4 and no size exit, scale = the largest bounding-box extent of
// the raw (uncentred) input.  Errors are printed, as there; the counts are
// then set to -1.
static int pamo_simplify_run(pamo_ctx* ctx, const float* in_verts, const int* in_tris, float* out_verts, int* out_tris,
                      int* in_out_nv_ni, float threshold) {
  const int nv = in_out_nv_ni[0], nt = in_out_nv_ni[1];
  if (nv < 0 || nt < 0) { ctx->fail("negative sizes"); return -1; }
  float bbox[6] = {-1e30f, -1e30f, -1e30f, 1e30f, 1e30f, 1e30f};
  for (int i = 0; i < nv; i++)
    for (int a = 0; a < 3; a++) {
      bbox[a] = std::max(bbox[a], in_verts[3 * i + a]);
      bbox[3 + a] = std::min(bbox[3 + a], in_verts[3 * i + a]);
    }
  const float scale = std::max({bbox[0] - bbox[3], bbox[1] - bbox[4], bbox[2] - bbox[5]});
  pamo_ctx& c = *ctx;
  if (!c.hip(hipSetDevice(c.device), "hipSetDevice") || !c.ensure(c.RV[1], (size_t)std::max(nv, 1) * 12) ||
      !c.ensure(c.RF[1], (size_t)std::max(nt, 1) * 12))
    return -1;
  c.st = nullptr;
  if (nv && !c.hip(hipMemcpy(c.RV[1].p, in_verts, (size_t)nv * 12, hipMemcpyHostToDevice), "copy in")) return -1;
  if (nt && !c.hip(hipMemcpy(c.RF[1].p, in_tris, (size_t)nt * 12, hipMemcpyHostToDevice), "copy in")) return -1;
  int oV = 0, oF = 0;
  int rc = pamo_run_stage2(ctx, nullptr, c.ptr<float>(c.RV[1]), c.ptr<int>(c.RF[1]), nv, nt, scale, -1, threshold, 4,
                           1000000, &oV, &oF, nullptr);
  if (rc) return rc;
  if (!c.ensure(c.RV[0], (size_t)std::max(oV, 1) * 12) || !c.ensure(c.RF[0], (size_t)std::max(oF, 1) * 12)) return -1;
  rc = pamo_run_fetch(ctx, nullptr, c.ptr<float>(c.RV[0]), c.ptr<int>(c.RF[0]));
  if (rc) return rc;
  if (oV && !c.hip(hipMemcpy(out_verts, c.RV[0].p, (size_t)oV * 12, hipMemcpyDeviceToHost), "copy out")) return -1;
  if (oF && !c.hip(hipMemcpy(out_tris, c.RF[0].p, (size_t)oF * 12, hipMemcpyDeviceToHost), "copy out")) return -1;
  in_out_nv_ni[0] = oV;
  in_out_nv_ni[1] = oF;
  return 0;
}

int pamo_prepare_host(const float* points, const int* faces, int V, int F, float* verts_out, float* tris_mean,
                      float* scale) {
  if (F <= 0 || V <= 0) return -1;
  // numpy: tris.mean(axis=1) is (((0 + a) + b) + c) / 3 per face and
  // coordinate; .mean(axis=0) adds the rows in order to 0, then divides by F
  // (add.reduce starts from its identity +0, which only matters for the sign
  // of a zero; numpy divides in double and rounds, for F < 2^24 the same as
  // a float division)
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (int f = 0; f < F; f++) {
    const float* a = points + 3LL * faces[3LL * f];
    const float* b = points + 3LL * faces[3LL * f + 1];
    const float* d = points + 3LL * faces[3LL * f + 2];
    for (int c = 0; c < 3; c++) acc[c] += (float)((double)(((0.0f + a[c]) + b[c]) + d[c]) / 3.0);
  }
  for (int c = 0; c < 3; c++) tris_mean[c] = (float)((double)acc[c] / (double)F);
  float lo[3], hi[3];
  for (long long i = 0; i < 3LL * V; i++) {
    const int c = (int)(i % 3);
    const float x = points[i] - tris_mean[c];
    verts_out[i] = x;
    if (i < 3) { lo[c] = hi[c] = x; continue; }
    // numpy's max and min propagate NaN
    if (x != x || x < lo[c]) lo[c] = (lo[c] != lo[c]) ? lo[c] : x;
    if (x != x || x > hi[c]) hi[c] = (hi[c] != hi[c]) ? hi[c] : x;
  }
  const float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
  // Python's max(max(e0, e1), e2)
  float m01 = e1 > e0 ? e1 : e0;
  *scale = e2 > m01 ? e2 : m01;
  return 0;
}

void pamo_simplify(float* in_verts, int* in_tris, float* out_verts, int* out_tris, int* in_out_nv_ni, float threshold) {
  static std::mutex mu;
  static pamo_ctx* ctx = nullptr;
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  if (!ctx && (hipGetDevice(&dev) != hipSuccess || pamo_ctx_create(&ctx, dev, PAMO_SHUFFLE_FEISTEL) != 0)) {
    fprintf(stderr, "pamo_simplify: no usable GPU\n");
    ctx = nullptr;
    in_out_nv_ni[0] = in_out_nv_ni[1] = -1;
    return;
  }
  if (pamo_simplify_run(ctx, in_verts, in_tris, out_verts, out_tris, in_out_nv_ni, threshold) != 0) {
    fprintf(stderr, "pamo_simplify: %s\n", pamo_last_error(ctx));
    in_out_nv_ni[0] = in_out_nv_ni[1] = -1;
  }
}

int pamo_set_strict(pamo_ctx* ctx, int strict) {
  if (!ctx) return -1;
  ctx->strict = strict ? 1 : 0;
  return 0;
}

int pamo_debug_guard(pamo_ctx* ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return -1;
  ctx->guard_mode = mode;
  return 0;
}

int pamo_set_tuning(pamo_ctx* ctx, const char* name, long long value) {
  if (!ctx || !name) return -1;
  if (!tuning_set(ctx->tune, name, value)) { ctx->fail("unknown tuning name or value out of range: %s", name); return -1; }
  return 0;
}

int pamo_get_tuning(pamo_ctx* ctx, const char* name, long long* value) {
  if (!ctx || !name || !value) return -1;
  for (const TuningField& f : kTuningFields)
    if (strcmp(f.name, name) == 0) { *value = ctx->tune.*(f.field); return 0; }
  return -1;
}

int pamo_set_placement(pamo_ctx* ctx, int mode) {
  if (!ctx || mode < 0 || mode > 1) return -1;
  ctx->placement = mode;
  return 0;
}

int pamo_debug_walk_split(pamo_ctx* ctx, int kshift) {
  if (!ctx || kshift < -1 || kshift > 3) return -1;
  ctx->walk_split = kshift;
  return 0;
}

int pamo_debug_pair_cap(pamo_ctx* ctx, long long cap) {
  if (!ctx) return -1;
  ctx->pair_cap_limit = cap;
  return 0;
}

int pamo_batch_state(pamo_ctx* ctx, long long out[6]) {
  if (!ctx) return -1;
  long long v[6] = {0, 0, 0, 0, 0, 0};
  if (ctx->eng) {
    Engine& g = *ctx->eng;
    v[0] = g.M;
    v[1] = g.V;
    v[2] = g.F;
    v[3] = g.st_checks;
    v[4] = g.st_queries;
    v[5] = g.st_cand;
  }
  memcpy(out, v, sizeof(v));
  return 0;
}

int pamo_ctx_create(pamo_ctx** out, int device, int shuffle_mode) {
  if (!out) return -1;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -3;
  if (device < 0 || device >= n) return -1;
  if (hipSetDevice(device) != hipSuccess) return -3;
  pamo_ctx* c = new pamo_ctx();
  c->device = device;
  c->shuffle_mode = shuffle_mode;
  if (const char* tf = getenv("PAMO_TUNING_FILE")) tuning_load_file(c->tune, tf);
  if (hipHostMalloc((void**)&c->hc, C_COUNT * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    delete c;
    return -3;
  }
  *out = c;
  return 0;
}

void pamo_ctx_destroy(pamo_ctx* ctx) { delete ctx; }

const char* pamo_last_error(const pamo_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

long long pamo_device_bytes(pamo_ctx* ctx) {
  if (!ctx) return -1;
  if (ctx->arena) return (long long)ctx->arena_cap;
  long long n = 0;
  for (Buf* b : ctx->bufs())
    if (b->p) n += (long long)b->bytes;
  if (ctx->eng) {
    for (Buf* b : eng_bufs(*ctx->eng))
      if (b->p) n += (long long)b->bytes;
    for (const EngSlot& s : ctx->eng->slots) n += (long long)(s.outV.p ? s.outV.bytes : 0) + (s.outF.p ? s.outF.bytes : 0);
  }
  if (ctx->s3)
    for (Buf* b : ctx->s3->buffers())
      if (b->p) n += (long long)b->bytes;
  return n;
}

int pamo_free_forward(pamo_ctx* ctx, void* stream, const float* d_points, const int* d_tris, const int* d_undo,
                      int n_undo, int n_pts, int n_tris, float scale, float threshold, int is_stuck, int init) {
  if (!ctx) return -1;
  ctx->st = (hipStream_t)stream;
  return ctx->forward(d_points, d_tris, d_undo, n_undo, n_pts, n_tris, scale, threshold, is_stuck, init);
}

int pamo_sp_forward(pamo_ctx* ctx, void* stream, const float* d_points, const int* d_tris, int n_pts, int n_tris,
                    float scale, float threshold, int init) {
  if (!ctx) return -1;
  ctx->st = (hipStream_t)stream;
  ctx->run_handle = -1;
  return ctx->forward(d_points, d_tris, nullptr, 0, n_pts, n_tris, scale, threshold, 0, init, 1);
}

int pamo_free_outputs(pamo_ctx* ctx, float** d_points, int** d_tris, int** d_occ, int** d_map, int** d_undo,
                      int* n_pts, int* n_tris, int* n_undo) {
  if (!ctx) return -1;
  if (d_points) *d_points = ctx->ptr<float>(ctx->P);
  if (d_tris) *d_tris = ctx->ptr<int>(ctx->T);
  if (d_occ) *d_occ = ctx->ptr<int>(ctx->occ);
  if (d_map) *d_map = ctx->ptr<int>(ctx->pts_map);
  if (d_undo) *d_undo = ctx->ptr<int>(ctx->undo_list);
  if (n_pts) *n_pts = ctx->V;
  if (n_tris) *n_tris = ctx->F;
  if (n_undo) *n_undo = ctx->n_vertices_undo;
  return 0;
}

int pamo_free_copy_outputs(pamo_ctx* ctx, void* stream, float* d_points, int* d_tris, int* d_occ, int* d_map,
                           int* d_undo) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  struct { void* dst; void* src; size_t n; } cp[5] = {
      {d_points, c.P.p, (size_t)c.V * 12}, {d_tris, c.T.p, (size_t)c.F * 12}, {d_occ, c.occ.p, (size_t)c.V * 4},
      {d_map, c.pts_map.p, (size_t)c.V * 4}, {d_undo, c.undo_list.p, (size_t)c.n_vertices_undo * 4}};
  for (auto& x : cp)
    if (x.dst && x.n && !c.hip(hipMemcpyAsync(x.dst, x.src, x.n, hipMemcpyDeviceToDevice, c.st), "copy outputs"))
      return -1;
  return 0;
}

int pamo_run_stage2(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts, int n_faces,
                    float scale, long long target, float threshold, int tolerance, long long iters, int* out_V,
                    int* out_F, long long* stats) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  int h = pamo_batch_submit(ctx, stream, d_verts, d_faces, n_verts, n_faces, scale, target, threshold, tolerance,
                            iters, -1);
  if (h < 0) return h;
  const long long s0[4] = {c.eng->st_checks, c.eng->st_queries, c.eng->st_cand, c.eng->st_hits};
  c.st = (hipStream_t)stream;
  if (!eng_rounds(c, INT_MAX, h, nullptr)) return -1;
  if (!c.eng->slots[h].done) { c.fail("engine stopped before the job finished"); return -1; }
  const EngSlot& s = c.eng->slots[h];
  c.run_faces = s.round_faces;
  c.run_handle = h;
  c.run_V = s.rV;
  c.run_F = s.rF;
  if (out_V) *out_V = s.rV;
  if (out_F) *out_F = s.rF;
  if (stats) {
    Engine& g = *c.eng;
    long long v[8] = {s.rounds,           s.exit_reason,          s.collapses,          s.max_passes,
                      g.st_checks - s0[0], g.st_queries - s0[1], g.st_cand - s0[2], g.st_hits - s0[3]};
    memcpy(stats, v, sizeof(v));
  }
  return 0;
}

int pamo_run_stage2_rounds(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts,
                           int n_faces, float scale, long long target, float threshold, int tolerance, long long iters,
                           int* out_V, int* out_F, long long* stats) {
  if (!ctx) return -1;
  ctx->run_handle = -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  c.err.clear();
  if (n_verts < 0 || n_faces < 0) { c.fail("negative size"); return -1; }
  if (!c.hip(hipSetDevice(c.device), "hipSetDevice")) return -1;
  for (int b = 0; b < 2; b++)
    if (!c.ensure(c.RV[b], (size_t)std::max(n_verts, 1) * 12) || !c.ensure(c.RF[b], (size_t)std::max(n_faces, 1) * 12))
      return -1;
  if (!c.ensure(c.keep, ((size_t)n_faces + 1) * 4) || !c.ensure(c.kpos, ((size_t)n_faces + 1) * 4) ||
      !c.ensure(c.undo_in, 256))
    return -1;
  int cur = 0;
  if (n_verts && !c.hip(hipMemcpyAsync(c.RV[0].p, d_verts, (size_t)n_verts * 12, hipMemcpyDeviceToDevice, c.st), "copy in"))
    return -1;
  if (n_faces && !c.hip(hipMemcpyAsync(c.RF[0].p, d_faces, (size_t)n_faces * 12, hipMemcpyDeviceToDevice, c.st), "copy in"))
    return -1;
  int V = n_verts, F = n_faces, n_undo = 0;
  int count = 0, is_stuck = 0;
  bool init = true;
  long long st[8] = {0, 4, 0, 0, 0, 0, 0, 0};
  c.run_faces.clear();
  for (long long it = 0; it < iters; it++) {
    int nprev = F;
    int rc = c.forward(c.ptr<float>(c.RV[cur]), c.ptr<int>(c.RF[cur]), c.ptr<int>(c.undo_in), n_undo, V, F, scale,
                       threshold, is_stuck, init ? 1 : 0);
    if (rc) return rc;
    init = false;
    st[0]++;
    st[2] += c.n_winners;
    st[3] = std::max<long long>(st[3], c.undo_passes);
    st[4] += c.n_checks;
    st[5] += c.n_queries;
    st[6] += c.n_cand;
    st[7] += c.n_hitfaces;
    // the undo list feeds the next round
    n_undo = c.n_vertices_undo;
    if (!c.ensure(c.undo_in, ((size_t)n_undo + 1) * 4)) return -1;
    if (n_undo && !c.hip(hipMemcpyAsync(c.undo_in.p, c.undo_list.p, (size_t)n_undo * 4, hipMemcpyDeviceToDevice, c.st),
                         "copy undo"))
      return -1;
    // compaction (simp_cuda/pamo/__init__.py:136-140)
    int nxt = cur ^ 1;
    k_compact_verts<<<nblk(c.V), BLOCK, 0, c.st>>>(c.ptr<float>(c.P), c.ptr<int>(c.occ), c.ptr<int>(c.pts_map), c.V,
                                                  c.ptr<float>(c.RV[nxt]));
    k_face_keep<<<nblk(c.F), BLOCK, 0, c.st>>>(c.ptr<int>(c.T), c.F, c.ptr<int>(c.keep));
    if (!c.check_launch("compact") || !c.scan(c.ptr<int>(c.keep), c.ptr<int>(c.kpos), c.F, C_NEWF)) return -1;
    k_compact_faces<<<nblk(c.F), BLOCK, 0, c.st>>>(c.ptr<int>(c.T), c.ptr<int>(c.keep), c.ptr<int>(c.kpos),
                                                  c.ptr<int>(c.pts_map), c.F, c.ptr<int>(c.RF[nxt]));
    if (!c.check_launch("compact faces") || !c.read_counters()) return -1;
    V = c.hc[C_NEWV];
    F = c.hc[C_NEWF];
    c.run_faces.push_back(F);
    cur = nxt;
    if (target >= 0 && (F <= target || F <= 10)) { st[1] = F <= target ? 1 : 2; break; }
    if (F == nprev) count++;
    else { count = 0; is_stuck = 0; }
    if (count >= 2) is_stuck = 1;
    if (count == tolerance) { st[1] = 3; break; }
  }
  c.run_cur = cur;
  c.run_V = V;
  c.run_F = F;
  if (out_V) *out_V = V;
  if (out_F) *out_F = F;
  if (stats) memcpy(stats, st, sizeof(st));
  return 0;
}

int pamo_run_rounds(pamo_ctx* ctx, int* host_faces, int cap) {
  if (!ctx) return -1;
  int n = (int)ctx->run_faces.size();
  if (host_faces)
    for (int i = 0; i < n && i < cap; i++) host_faces[i] = ctx->run_faces[i];
  return n;
}

int pamo_run_fetch(pamo_ctx* ctx, void* stream, float* d_verts_out, int* d_faces_out) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  if (c.run_handle >= 0) return pamo_batch_take(ctx, stream, c.run_handle, d_verts_out, d_faces_out);
  if (c.run_V && d_verts_out &&
      !c.hip(hipMemcpyAsync(d_verts_out, c.RV[c.run_cur].p, (size_t)c.run_V * 12, hipMemcpyDeviceToDevice, c.st), "fetch"))
    return -1;
  if (c.run_F && d_faces_out &&
      !c.hip(hipMemcpyAsync(d_faces_out, c.RF[c.run_cur].p, (size_t)c.run_F * 12, hipMemcpyDeviceToDevice, c.st), "fetch"))
    return -1;
  return c.hip(hipStreamSynchronize(c.st), "sync") ? 0 : -1;
}

// ------------------------------------------------ no-allocation variant
// Device region of pamo_run_stage2_noalloc: every engine buffer of a call
// is carved from it.  Each GPU the library is loaded on holds its own copy
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (na_region), never
// through the array's name.
static __device__ __align__(256) unsigned char scratch[PAMO_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
struct NoAllocDevice {
  std::mutex mu;            // one call at a time per device (in this process)
  char* region = nullptr;   // this device's address of scratch
  pamo_ctx* c = nullptr;    // host state of the last call (its buffers point into the region)
};
NoAllocDevice na_dev[NA_MAX_DEVICES];
}  // namespace

int pamo_run_stage2_noalloc(pamo_ctx* ctx, void* stream, const float* d_verts, const int* d_faces, int n_verts,
                            int n_faces, float scale, long long target, float threshold, int tolerance,
                            long long iters, float* d_verts_out, int* d_faces_out, int* out_V, int* out_F,
                            long long* stats) {
  if (!ctx) return -1;
  ctx->err.clear();
  if (n_verts < 0 || n_faces < 0 || n_verts > PAMO_NOALLOC_MAX_VERTS || n_faces > PAMO_NOALLOC_MAX_FACES) {
    ctx->fail("pamo_run_stage2_noalloc takes at most %d vertices and %d faces (got %d and %d)",
              PAMO_NOALLOC_MAX_VERTS, PAMO_NOALLOC_MAX_FACES, n_verts, n_faces);
    return -4;
  }
  if (ctx->strict) {
    ctx->fail("pamo_run_stage2_noalloc runs the reference guard only (strict mode is set on this context)");
    return -4;
  }
  if ((n_verts && !d_verts_out) || (n_faces && !d_faces_out)) {
    ctx->fail("pamo_run_stage2_noalloc: output buffers are required");
    return -1;
  }
  if (ctx->device < 0 || ctx->device >= NA_MAX_DEVICES) { ctx->fail("device %d out of range", ctx->device); return -1; }
  NoAllocDevice& nd = na_dev[ctx->device];
  std::lock_guard<std::mutex> lock(nd.mu);
  if (!ctx->hip(hipSetDevice(ctx->device), "hipSetDevice")) return -1;
  if (!nd.region) {   // set-up, once per device, with that device current
    void* p = nullptr;
    if (!ctx->hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return -1;
    nd.region = (char*)p;
  }
  // fresh host state over the same region: every buffer is carved again
  // from the region's start, and nothing of the last call is read before
  // this call writes it (counters and scan descriptors are cleared by
  // kernels, the engine's persistent regions are zero-filled at admission)
  delete nd.c;
  nd.c = new pamo_ctx();
  pamo_ctx& c = *nd.c;
  c.device = ctx->device;
  c.shuffle_mode = ctx->shuffle_mode;
  c.guard_mode = ctx->guard_mode;
  c.pair_cap_limit = ctx->pair_cap_limit;
  c.walk_split = ctx->walk_split;
  c.placement = ctx->placement;
  c.tune = ctx->tune;
  c.arena = nd.region;
  c.arena_cap = PAMO_NOALLOC_SCRATCH_BYTES;
  c.hc = c.hc_host;
  int oV = 0, oF = 0;
  int rc = pamo_run_stage2(&c, stream, d_verts, d_faces, n_verts, n_faces, scale, target, threshold, tolerance, iters,
                           &oV, &oF, stats);
  if (rc == 0) rc = pamo_run_fetch(&c, stream, d_verts_out, d_faces_out);
  if (rc != 0) {
    ctx->err = c.err;
    return rc;
  }
  if (out_V) *out_V = oV;
  if (out_F) *out_F = oF;
  return 0;
}

int pamo_profile(pamo_ctx* ctx, unsigned kernel_mask) {
  if (!ctx) return -1;
  ctx->prof_resolve();
  ctx->prof_mask = kernel_mask;
  for (int k = 0; k < 32; k++) { ctx->prof_ms[k] = 0; ctx->prof_bytes[k] = 0; ctx->prof_n[k] = 0; }
  return 0;
}

int pamo_profile_read(pamo_ctx* ctx, int kernel, double out[3]) {
  if (!ctx || kernel < 0 || kernel >= 32) return -1;
  ctx->prof_resolve();
  out[0] = (double)ctx->prof_n[kernel];
  out[1] = ctx->prof_ms[kernel];
  out[2] = ctx->prof_bytes[kernel];
  return 0;
}

long long pamo_debug_copy(pamo_ctx* ctx, void* stream, int which, void* host_dst, long long cap_bytes) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  const void* src = nullptr;
  long long n = 0;
  switch (which) {
    case 5: src = c.fan_off.p; n = 4LL * (c.V + 1); break;
    case 6: src = c.fan.p; n = 4LL * c.n_fan; break;
    case 7: src = c.edges.p; n = 8LL * c.E; break;
    case 8: n = 64LL * c.V; break;
    case 9: src = c.edge_cost.p; n = 4LL * c.E; break;
    case 10: src = c.tri_min.p; n = 8LL * c.F; break;
    case 11: src = c.collapsed.p; n = 4LL * c.n_collapsed; break;
    case 12: src = c.have_s_first ? c.S_first.p : nullptr; n = c.F; break;
    case 13: src = c.T0.p; n = 12LL * c.F; break;
    case 14: src = c.table.p; n = c.alloc_pts ? (long long)c.alloc_pts + 1 : 0; break;
    default: c.fail("unknown debug buffer %d", which); return -1;
  }
  if (!host_dst || n > cap_bytes || n == 0) return n;
  if (which == 8) {
    std::vector<float> q((size_t)10 * c.V);
    if (!c.hip(hipMemcpyAsync(q.data(), c.Q.p, q.size() * 4, hipMemcpyDeviceToHost, c.st), "copy") ||
        !c.hip(hipStreamSynchronize(c.st), "sync"))
      return -1;
    static const int map16[16] = {0, 1, 2, 3, 1, 4, 5, 6, 2, 5, 7, 8, 3, 6, 8, 9};
    float* o = (float*)host_dst;
    for (int v = 0; v < c.V; v++)
      for (int r = 0; r < 16; r++) o[16LL * v + r] = q[10LL * v + map16[r]];
    return n;
  }
  if (!src) {
    memset(host_dst, 0, (size_t)n);
    return n;
  }
  if (!c.hip(hipMemcpyAsync(host_dst, src, (size_t)n, hipMemcpyDeviceToHost, c.st), "copy") ||
      !c.hip(hipStreamSynchronize(c.st), "sync"))
    return -1;
  return n;
}

long long pamo_debug_pairs(pamo_ctx* ctx, void* stream, const float* d_points, const int* d_tris, int n_pts,
                           int n_tris, int* host_pairs, long long cap_rows) {
  if (!ctx) return -1;
  pamo_ctx& c = *ctx;
  c.st = (hipStream_t)stream;
  c.err.clear();
  if (!c.hip(hipSetDevice(c.device), "hipSetDevice")) return -1;
  const long long Vl = std::max(n_pts, 1), Fl = std::max(n_tris, 1);
  if (!c.ensure(c.counters, C_COUNT * 4) || !c.ensure(c.cstats, 64 * 4) || !c.ensure(c.P, Vl * 12) ||
      !c.ensure(c.T, Fl * 12) || !c.ensure(c.af, Fl * 4) || !c.ensure(c.apos, Fl * 4) || !c.ensure(c.qlist, Fl * 4) ||
      !c.ensure(c.leaf_face, Fl * 4) || !c.ensure(c.S, Fl))
    return -1;
  c.V = n_pts;
  c.F = n_tris;
  if (!c.fill_i32(c.counters.p, 0, C_COUNT)) return -1;
  k_validate<<<nblk(n_tris), BLOCK, 0, c.st>>>(d_tris, n_tris, n_pts, c.cnt_dev(C_ERR));
  if (!c.hip(hipMemcpyAsync(c.P.p, d_points, (size_t)n_pts * 12, hipMemcpyDeviceToDevice, c.st), "copy") ||
      !c.hip(hipMemcpyAsync(c.T.p, d_tris, (size_t)n_tris * 12, hipMemcpyDeviceToDevice, c.st), "copy"))
    return -1;
  if (!c.read_counters()) return -1;
  if (c.hc[C_ERR]) { c.fail("invalid faces"); return -2; }
  // all alive faces are both leaves and queries
  if (!c.ensure(c.qflag, Fl) || !c.fill_u8(c.qflag.p, 1, Fl) || !c.ensure(c.qf, Fl * 4) || !c.ensure(c.qpos, (Fl + 1) * 4))
    return -1;
  k_query_flags<<<nblk(n_tris), BLOCK, 0, c.st>>>(c.ptr<uint8_t>(c.qflag), c.ptr<int>(c.T), n_tris, c.ptr<int>(c.qf),
                                                 c.ptr<int>(c.af));
  if (!c.scan(c.ptr<int>(c.af), c.ptr<int>(c.apos), n_tris, C_NALIVE)) return -1;
  k_compact_idx<<<nblk(n_tris), BLOCK, 0, c.st>>>(c.ptr<int>(c.af), c.ptr<int>(c.apos), n_tris, c.ptr<int>(c.leaf_face), 0);
  if (!c.scan(c.ptr<int>(c.qf), c.ptr<int>(c.qpos), n_tris, C_NQUERY)) return -1;
  k_compact_idx<<<nblk(n_tris), BLOCK, 0, c.st>>>(c.ptr<int>(c.qf), c.ptr<int>(c.qpos), n_tris, c.ptr<int>(c.qlist), 0);
  std::vector<int> pairs;
  if (!c.build_bvh_and_traverse(n_tris, true, &pairs)) return -1;
  long long np = (long long)pairs.size() / 3;
  if (host_pairs && np <= cap_rows) memcpy(host_pairs, pairs.data(), pairs.size() * 4);
  return np;
}

}  // extern "C"
