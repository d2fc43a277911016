"""ctypes bindings of the CPU oracle and of the exact mesh verifier.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's ``cpu_baseline`` leg, always as the checker.  The product
package ``pamo_amd`` never imports this module.

The oracle restates the reference's stage-2 round
(simp_cuda/src/cusimp_free.cu:891-1048) and the ``PaMO.run`` loop
(simp_cuda/pamo/__init__.py:95-178); see pamo_oracle.cpp for the arithmetic
spec and the parity-mode definitions D0-D9.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PAMO_ORACLE_LIB names another build of this oracle (the mutation tests load
# deliberately broken copies through it); by default the one built here
_ORC = os.environ.get("PAMO_ORACLE_LIB") or os.path.join(_HERE, "libpamo_oracle.so")

SHUFFLE_IDENTITY = 0
SHUFFLE_FEISTEL = 1


def build():
    """Compile the oracle (make in this directory)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load(path):
    if not os.path.exists(path):
        if os.environ.get("PAMO_ORACLE_LIB"):
            raise FileNotFoundError(path)
        build()
    return ctypes.CDLL(path)


_orc = None


def _lib():
    global _orc
    if _orc is None:
        _orc = _load(_ORC)
        P = ctypes.c_void_p
        _orc.orc_create.restype = P
        _orc.orc_destroy.argtypes = [P]
        _orc.orc_forward.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _orc.orc_forward_sp.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_int, ctypes.c_int]
        _orc.orc_sizes.argtypes = [P, P]
        _orc.orc_copy.argtypes = [P, ctypes.c_int, P, ctypes.c_longlong]
        _orc.orc_copy.restype = ctypes.c_longlong
        _orc.orc_run.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_longlong, ctypes.c_float,
                                 ctypes.c_int, ctypes.c_longlong, ctypes.c_int, P, P, P, P, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int]
        _orc.orc_run_shadow.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_longlong,
                                        ctypes.c_float, ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                        ctypes.c_double, P]
        _orc.orc_place.argtypes = [P, P, P, P]
        _orc.orc_place.restype = ctypes.c_int
        _orc.orc_shuffle_perm.argtypes = [ctypes.c_longlong, P]
        _orc.orc_narrow.argtypes = [P, P, P, P]
        _orc.orc_isect_raw.argtypes = [P, P, P, P]
        _orc.orc_is_coplanar.argtypes = [P, P]
        _orc.orc_full_check.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_longlong, P, ctypes.c_longlong, P]
        _orc.orc_set_threads.argtypes = [ctypes.c_int]
        _orc.orc_run_timed.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_longlong,
                                       ctypes.c_float, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_double,
                                       P, P, P, ctypes.c_int]
        _orc.orc_cost_twin.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, P, P, P, P,
                                       P, P, ctypes.c_int]
        _orc.orc_mesh_to_sdf.argtypes = [P, ctypes.c_longlong, ctypes.c_int, ctypes.c_float, P]
        _orc.orc_volume_to_mesh.argtypes = [P, ctypes.c_int, P]
        _orc.orc_mesh_fetch.argtypes = [P, P]
    return _orc


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def set_threads(n):
    """Set the oracle's OpenMP thread count; returns the count in effect."""
    return _lib().orc_set_threads(int(n))


BUF = dict(points=0, tris=1, occ=2, pts_map=3, undo=4, fan_off=5, fan=6, edges=7, Q=8, edge_cost=9,
           tri_min=10, collapsed=11, S=12, orig_tris=13, table=14)
_DT = dict(points=np.float32, tris=np.int32, occ=np.int32, pts_map=np.int32, undo=np.int32, fan_off=np.int32,
           fan=np.int32, edges=np.int32, Q=np.float32, edge_cost=np.uint32, tri_min=np.uint64, collapsed=np.int32,
           S=np.uint8, orig_tris=np.int32, table=np.uint8)
SIZE_KEYS = ["V", "F", "E", "n_vertices_undo", "n_collapsed", "n_winners", "undo_passes", "n_checks", "n_fan",
             "undo_cut", "collapse_t", "alloc_pts", "n_queries", "n_candidates", "n_hits"]


class Oracle:
    """Stateful CPU restatement of ``CUSimp_Free`` (one object per context)."""

    def __init__(self, shuffle=SHUFFLE_FEISTEL):
        self._lib = _lib()
        self._ctx = self._lib.orc_create()
        self.shuffle = shuffle

    def __del__(self):
        try:
            self._lib.orc_destroy(self._ctx)
        except Exception:
            pass

    def forward(self, points, tris, undo, scale, threshold, is_stuck, init):
        points = np.ascontiguousarray(points, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        undo = np.ascontiguousarray(undo, np.int32)
        rc = self._lib.orc_forward(self._ctx, _ptr(points), _ptr(tris), _ptr(undo) if undo.size else None,
                                   int(undo.size), int(points.shape[0]), int(tris.shape[0]), float(scale),
                                   float(threshold), int(bool(is_stuck)), int(bool(init)), int(self.shuffle))
        if rc != 0:
            raise RuntimeError("oracle forward failed with code %d" % rc)
        return self.outputs()

    def forward_sp(self, points, tris, scale, threshold, init):
        """The unguarded round of ``CUDSP.forward`` (simp_cuda/src/cusimp.cu:
        587-720); returns (points, tris, occ, map)."""
        points = np.ascontiguousarray(points, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        rc = self._lib.orc_forward_sp(self._ctx, _ptr(points), _ptr(tris), int(points.shape[0]), int(tris.shape[0]),
                                      float(scale), float(threshold), int(bool(init)), int(self.shuffle))
        if rc != 0:
            raise RuntimeError("oracle forward_sp failed with code %d" % rc)
        return (self.get("points"), self.get("tris"), self.get("occ").reshape(-1, 1),
                self.get("pts_map").reshape(-1, 1))

    def sizes(self):
        out = np.zeros(16, np.int64)
        self._lib.orc_sizes(self._ctx, _ptr(out))
        return dict(zip(SIZE_KEYS, [int(x) for x in out[:len(SIZE_KEYS)]]))

    def get(self, name):
        which = BUF[name]
        n = self._lib.orc_copy(self._ctx, which, None, 0)
        dt = np.dtype(_DT[name])
        a = np.empty(n // dt.itemsize, dt)
        self._lib.orc_copy(self._ctx, which, _ptr(a), n)
        if name in ("points", "tris", "orig_tris"):
            a = a.reshape(-1, 3)
        elif name == "edges":
            a = a.reshape(-1, 2)
        elif name == "Q":
            a = a.reshape(-1, 16)
        return a

    def outputs(self):
        """The five outputs of ``CUDSP_Free.forward`` (pybind.cpp:74-102)."""
        return (self.get("points"), self.get("tris"), self.get("occ").reshape(-1, 1),
                self.get("pts_map").reshape(-1, 1), self.get("undo"))


def run(verts, faces, scale, target, threshold=1e-3, tolerance=4, iters=1000000, shuffle=SHUFFLE_FEISTEL,
        return_rounds=False, strict=False, placement=0):
    """The stage-2 round loop of ``PaMO.run`` on centred vertices (C++).
    strict: the strict guard (not in the reference): exact pair test added
    to the reference's, whole collapsed list scanned, undo until clean.
    placement: 0 the edge midpoint (the reference's), 1 the optimal vertex
    (not in the reference; pamo_amd/csrc/placement.h)."""
    lib = _lib()
    verts = np.ascontiguousarray(verts, np.float32)
    faces = np.ascontiguousarray(faces, np.int32)
    V, F = verts.shape[0], faces.shape[0]
    ov = np.empty((V, 3), np.float32)
    of = np.empty((F, 3), np.int32)
    st = np.zeros(8, np.int64)
    cap = 100000
    rf = np.zeros(cap, np.int32)
    rc = lib.orc_run(_ptr(verts), _ptr(faces), V, F, float(scale), int(target), float(threshold), int(tolerance),
                     int(iters), int(shuffle), _ptr(ov), _ptr(of), _ptr(st), _ptr(rf), cap, int(bool(strict)),
                     int(placement))
    if rc != 0:
        raise RuntimeError("oracle run failed with code %d" % rc)
    stats = dict(rounds=int(st[0]), V=int(st[1]), F=int(st[2]), exit_reason=int(st[3]),
                 total_collapsed=int(st[4]), max_undo_passes=int(st[6]))
    out = (ov[:stats["V"]].copy(), of[:stats["F"]].copy(), stats)
    if return_rounds:
        return out + (rf[:stats["rounds"]].copy(),)
    return out


def run_timed(verts, faces, scale, target, max_seconds, threshold=1e-3, tolerance=4, shuffle=SHUFFLE_FEISTEL):
    """The same loop stopped after the first round that ends past
    max_seconds.  Returns (stats, faces after each round, seconds elapsed at
    the end of each round)."""
    lib = _lib()
    verts = np.ascontiguousarray(verts, np.float32)
    faces = np.ascontiguousarray(faces, np.int32)
    st = np.zeros(8, np.int64)
    cap = 100000
    rf = np.zeros(cap, np.int32)
    rs = np.zeros(cap, np.float64)
    rc = lib.orc_run_timed(_ptr(verts), _ptr(faces), verts.shape[0], faces.shape[0], float(scale), int(target),
                           float(threshold), int(tolerance), 1000000, int(shuffle), float(max_seconds), _ptr(st),
                           _ptr(rf), _ptr(rs), cap)
    if rc != 0:
        raise RuntimeError("oracle run failed with code %d" % rc)
    n = int(st[0])
    return dict(rounds=n, exit_reason=int(st[3])), rf[:n].copy(), rs[:n].copy()


def run_rounds(verts, faces, scale, target, threshold=1e-3, tolerance=4, iters=1000000, shuffle=SHUFFLE_FEISTEL,
               progress=None):
    """``run`` as the reference's Python loop (simp_cuda/pamo/__init__.py:
    122-158): one ``Oracle.forward`` per round, the compaction by occ and map
    in numpy, and progress(rounds, faces) after every round (for long runs
    that report as they go).  Returns (verts, faces, stats, faces after each
    round) like ``run(..., return_rounds=True)``."""
    o = Oracle(shuffle)
    pts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(faces, np.int32)
    undo = np.zeros(0, np.int32)
    count, is_stuck, init, rounds, exit_reason, rf = 0, 0, True, 0, 4, []
    for _ in range(iters):
        n_prev = tris.shape[0]
        p2, t2, occ, vmap, undo = o.forward(pts, tris, undo, scale, threshold, is_stuck, init)
        init = False
        rounds += 1
        pts = p2[occ.reshape(-1).astype(bool)]
        t2 = t2[t2[:, 0] >= 0]
        tris = np.ascontiguousarray(vmap.reshape(-1)[t2], np.int32)
        n_cur = tris.shape[0]
        rf.append(n_cur)
        if progress is not None:
            progress(rounds, n_cur)
        if n_cur <= target or n_cur <= 10:
            exit_reason = 1 if n_cur <= target else 2
            break
        if n_cur == n_prev:
            count += 1
        else:
            count = 0
            is_stuck = 0
        if count >= 2:
            is_stuck = 1
        if count == tolerance:
            exit_reason = 3
            break
    stats = dict(rounds=rounds, V=int(pts.shape[0]), F=int(tris.shape[0]), exit_reason=exit_reason)
    return pts, tris, stats, np.array(rf, np.int32)


def fp64_shadow(verts, faces, scale, target, threshold=1e-3, tolerance=4, max_seconds=0.0,
                shuffle=SHUFFLE_FEISTEL):
    """The FP32 round loop with an FP64 shadow of every round: costs of all
    edges from the FP64 twin on the round's start state, the winners those
    costs would choose, and double positions that follow the FP32 run's
    collapses (see orc_run_shadow).  max_seconds > 0 stops after the first
    round that ends past it (exit_reason 5)."""
    verts = np.ascontiguousarray(verts, np.float32)
    faces = np.ascontiguousarray(faces, np.int32)
    out = np.zeros(15, np.float64)
    rc = _lib().orc_run_shadow(_ptr(verts), _ptr(faces), verts.shape[0], faces.shape[0], float(scale), int(target),
                               float(threshold), int(tolerance), 1000000, int(shuffle), float(max_seconds),
                               _ptr(out))
    if rc != 0:
        raise RuntimeError("oracle shadow run failed with code %d" % rc)
    return dict(rounds=int(out[0]), edges_scored=int(out[1]), edges_valued=int(out[2]),
                quantized_cost_differs=int(out[3]), winners_fp32=int(out[4]), winners_fp64=int(out[5]),
                winners_differ=int(out[6]), cost_rel_err_median=float(out[7]), cost_rel_err_p99=float(out[8]),
                cost_rel_err_max=float(out[9]), position_err_max=float(out[10]), position_err_mean=float(out[11]),
                exit_reason=int(out[12]), final_faces=int(out[13]), seconds=round(float(out[14]), 2))


def place(q, a, b):
    """The optimal placement (pamo_amd/csrc/placement.h) of edge (a, b) under
    the quadric q (10 unique entries): (position, branch taken)."""
    q = np.ascontiguousarray(q, np.float32)
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(3, np.float32)
    how = _lib().orc_place(_ptr(q), _ptr(a), _ptr(b), _ptr(out))
    return out, int(how)


def shuffle_perm(m):
    """Output slot j of the init shuffle gathers input face perm[j]."""
    out = np.empty(max(m, 0), np.int64)
    _lib().orc_shuffle_perm(int(m), _ptr(out))
    return out


def narrow(cur, qry, ci, qi):
    cur = np.ascontiguousarray(cur, np.float32).reshape(9)
    qry = np.ascontiguousarray(qry, np.float32).reshape(9)
    ci = np.ascontiguousarray(ci, np.int32).reshape(3)
    qi = np.ascontiguousarray(qi, np.int32).reshape(3)
    return int(_lib().orc_narrow(_ptr(cur), _ptr(qry), _ptr(ci), _ptr(qi)))


def isect_raw(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(9)
    b = np.ascontiguousarray(b, np.float32).reshape(9)
    r1 = np.zeros(3, np.float32)
    r2 = np.zeros(3, np.float32)
    hit = _lib().orc_isect_raw(_ptr(a), _ptr(b), _ptr(r1), _ptr(r2))
    return bool(hit), r1, r2


def is_coplanar(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(9)
    b = np.ascontiguousarray(b, np.float32).reshape(9)
    return bool(_lib().orc_is_coplanar(_ptr(a), _ptr(b)))


def cost_twin(points, faces, scale, threshold=1e-3):
    """First-round edge costs of a fresh mesh (no shuffle) evaluated in
    float and in double from the same float positions: dict of c32, c64
    (unquantized values), st32, st64 (0 value, 1 maximum, 2 unchanged), q32
    (the float value quantized) and mid (relative midpoint difference)."""
    lib = _lib()
    pts = np.ascontiguousarray(points, np.float32)
    tri = np.ascontiguousarray(faces, np.int32)
    cap = 2 * len(tri) + 1
    out = dict(c32=np.zeros(cap, np.float32), c64=np.zeros(cap, np.float64), st32=np.zeros(cap, np.int32),
               st64=np.zeros(cap, np.int32), q32=np.zeros(cap, np.uint32), mid=np.zeros(cap, np.float64))
    E = lib.orc_cost_twin(_ptr(pts), _ptr(tri), len(pts), len(tri), float(scale), float(threshold),
                          _ptr(out["c32"]), _ptr(out["c64"]), _ptr(out["st32"]), _ptr(out["st64"]), _ptr(out["q32"]),
                          _ptr(out["mid"]), cap)
    if E < 0:
        raise RuntimeError("orc_cost_twin failed")
    return {k: v[:E] for k, v in out.items()}


def precision_report(points, faces, scale, threshold=1e-3):
    """FP32 against FP64 on the first round: the relative error of the
    unquantized cost (edges valued in both), the share of edges whose
    quantized cost or status differs, and the largest relative midpoint
    (collapse position) difference."""
    t = cost_twin(points, faces, scale, threshold)
    both = (t["st32"] == 0) & (t["st64"] == 0)
    c64 = t["c64"][both]
    rel = np.abs(t["c32"][both].astype(np.float64) - c64) / np.maximum(np.abs(c64), 1e-300)
    q64 = np.where(t["st64"] == 1, np.float64(4294967295.0),
                   np.floor(np.clip(t["c64"], 0.0, 10.0) / 10.0 * 4294967296.0))
    q64 = np.minimum(q64, 4294967295.0).astype(np.uint64)
    qdiff = (t["q32"].astype(np.uint64) != q64) & (t["st32"] != 2)
    # first-round winners under each cost (unchanged costs are 0 in a fresh
    # context): the claim of cusimp_free.cu:587-681 as vertex minima
    tri = np.ascontiguousarray(faces, np.int64)
    alive = ~(tri == -1).all(axis=1)
    a = tri[alive]
    parts = []
    for (x, y) in ((0, 1), (1, 2), (2, 0)):
        sel = a[:, x] > a[:, y]
        parts.append((np.where(sel)[0] * 3 + x, a[sel][:, [y, x]]))
    order = np.argsort(np.concatenate([p[0] for p in parts]), kind="stable")
    edges = np.concatenate([p[1] for p in parts])[order]
    ct = np.uint64(int(np.float32(min(max(threshold, 0.0), 10.0)) / np.float32(10.0) * np.float32(4294967296.0)))

    def winners(q):
        q = np.where(t["st32"] == 2, 0, q).astype(np.uint64)
        key = (q << np.uint64(32)) | np.arange(len(q), dtype=np.uint64)
        V = int(a.max()) + 1 if a.size else 0
        vmin = np.full(V, np.iinfo(np.uint64).max, np.uint64)
        np.minimum.at(vmin, edges[:, 0], key)
        np.minimum.at(vmin, edges[:, 1], key)
        fmin = np.minimum(np.minimum(vmin[a[:, 0]], vmin[a[:, 1]]), vmin[a[:, 2]])
        vf = np.full(V, np.iinfo(np.uint64).max, np.uint64)
        for k in range(3):
            np.minimum.at(vf, a[:, k], fmin)
        u, v = edges[:, 0], edges[:, 1]
        return (q <= ct) & (vmin[u] == key) & (vmin[v] == key) & (vf[u] == key) & (vf[v] == key)

    w32, w64 = winners(t["q32"]), winners(q64)
    return dict(edges=int(len(t["c32"])), valued=int(both.sum()), status_differs=int((t["st32"] != t["st64"]).sum()),
                cost_rel_err_median=float(np.median(rel)) if rel.size else 0.0,
                cost_rel_err_p99=float(np.percentile(rel, 99)) if rel.size else 0.0,
                cost_rel_err_max=float(rel.max()) if rel.size else 0.0,
                quantized_differs=int(qdiff.sum()),
                winners_fp32=int(w32.sum()), winners_fp64=int(w64.sum()), winners_differ=int((w32 != w64).sum()),
                position_rel_err_max=float(t["mid"].max()) if len(t["mid"]) else 0.0)


def full_check(points, tris):
    """All ordered candidate pairs (q, c) and intersecting pairs of a mesh."""
    points = np.ascontiguousarray(points, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    cnt = np.zeros(2, np.int64)
    _lib().orc_full_check(_ptr(points), _ptr(tris), points.shape[0], tris.shape[0], None, 0, None, 0, _ptr(cnt))
    cand = np.empty((int(cnt[0]), 2), np.int32)
    hits = np.empty((int(cnt[1]), 2), np.int32)
    _lib().orc_full_check(_ptr(points), _ptr(tris), points.shape[0], tris.shape[0], _ptr(cand), cand.size + 1,
                          _ptr(hits), hits.size + 1, _ptr(cnt))
    return cand, hits


def preprocess(points, faces):
    """Centring and scale of ``PaMO.run`` with stage 1 off
    (simp_cuda/pamo/__init__.py:60-64,101,117,126), in numpy float32."""
    points = np.asarray(points, np.float32)
    faces = np.asarray(faces)
    tris = points[faces]
    tris_mean = tris.mean(axis=1).mean(axis=0)
    verts = points - tris_mean
    ext = verts.max(axis=0) - verts.min(axis=0)
    scale = max(max(ext[0], ext[1]), ext[2])
    return verts.astype(np.float32), np.float32(scale), tris_mean


def mesh_to_sdf(tris, R, band):
    """Stage 1's signed distance lattice (R^3 float32, index [k, j, i]) of
    triangles given as [n, 3, 3] float32 in [0, 1]^3 (pamo_mesh_to_sdf's
    definition)."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    vol = np.empty((R, R, R), np.float32)
    rc = _lib().orc_mesh_to_sdf(_ptr(tris), tris.shape[0], int(R), float(band), _ptr(vol))
    if rc != 0:
        raise RuntimeError("orc_mesh_to_sdf failed (%d)" % rc)
    return vol


def volume_to_mesh(vol):
    """Stage 1's extraction of an R^3 volume's zero level set (positive
    outside): (verts [V, 3] float32 in [0, 1]^3, faces [F, 3] int32)."""
    vol = np.ascontiguousarray(vol, np.float32)
    R = vol.shape[0]
    cnt = np.zeros(2, np.int64)
    rc = _lib().orc_volume_to_mesh(_ptr(vol), int(R), _ptr(cnt))
    if rc != 0:
        raise RuntimeError("orc_volume_to_mesh failed (%d)" % rc)
    v = np.empty((int(cnt[0]), 3), np.float32)
    f = np.empty((int(cnt[1]), 3), np.int32)
    _lib().orc_mesh_fetch(_ptr(v), _ptr(f))
    return v, f
