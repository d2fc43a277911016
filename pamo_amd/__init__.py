"""MI355X-native PaMO stage 2: guarded parallel QEM edge collapse.

Drop-in for the reference's ``pamo`` package on its stage-2 path
(simp_cuda/pamo/__init__.py): ``PaMO(mesh, use_stage1, use_stage3).run(
points, triangles, ratio, tolerance, threshold, iter, min_verts)`` and the
extension module ``pamo_amd._C`` (``CUDSP_Free``).  The rounds run in
libpamo_hip.so (hand-written gfx950 HIP kernels), as does stage 1's
remeshing (``pamo_amd.remesh``) and stage 3's safe projection
(``pamo_amd.safe_project``).
"""
import time

import numpy as np
import torch

from . import _C, _lib
from ._lib import PamoError

__all__ = ["PaMO", "PaSP", "PamoError", "simplify_stage2", "preprocess_numpy", "prepare_host"]


def _as_vf(input_mesh):
    if isinstance(input_mesh, (tuple, list)) and len(input_mesh) == 2:
        return np.asarray(input_mesh[0]), np.asarray(input_mesh[1])
    return np.asarray(input_mesh.vertices), np.asarray(input_mesh.faces)


def _preprocess(points, triangles):
    """tris_mean exactly as simp_cuda/pamo/__init__.py:60-64 (numpy float32)."""
    tris = points[triangles.long()]
    tris = tris.cpu().numpy()
    return tris.mean(axis=1).mean(axis=0)


def preprocess_numpy(points, faces):
    """Centring and scale of PaMO.run (simp_cuda/pamo/__init__.py:60-64,117,126)
    on host arrays: returns (centred float32 vertices, scale, tris_mean)."""
    points = np.asarray(points, np.float32)
    tris_mean = points[np.asarray(faces)].mean(axis=1).mean(axis=0)
    verts = (points - tris_mean).astype(np.float32)
    ext = verts.max(axis=0) - verts.min(axis=0)
    return verts, float(max(max(ext[0], ext[1]), ext[2])), tris_mean


def prepare_host(points, faces):
    """PaMO.run's centring and scale (simp_cuda/pamo/__init__.py:60-64,117,
    126) on the host arrays, in the library (pamo_prepare_host: numpy's
    float32 summation order, one pass, no temporaries; the call releases the
    GIL, so worker threads run it in parallel).  Returns (centred float32
    vertices, int32 faces, scale, tris_mean)."""
    pts = np.ascontiguousarray(points, np.float32)
    tri = np.ascontiguousarray(faces, np.int32)
    if len(tri) == 0 or len(pts) == 0 or tri.min() < 0 or tri.max() >= len(pts):
        verts, scale, tris_mean = preprocess_numpy(pts, tri)   # numpy raises on bad indices
        return verts, tri, scale, tris_mean
    verts = np.empty_like(pts)
    mean = np.empty(3, np.float32)
    scale = np.empty(1, np.float32)
    rc = _lib.lib().pamo_prepare_host(pts.ctypes.data, tri.ctypes.data, len(pts), len(tri), verts.ctypes.data,
                                       mean.ctypes.data, scale.ctypes.data)
    if rc != 0:
        raise RuntimeError("pamo_prepare_host failed")
    return verts, tri, float(scale[0]), mean


def _scale(verts):
    """simp_cuda/pamo/__init__.py:126: the largest axis extent."""
    ext = [verts[:, k].max() - verts[:, k].min() for k in range(3)]
    return float(max(max(ext[0], ext[1]), ext[2]))


def simplify_stage2(verts, faces, scale, target, threshold=1e-3, tolerance=4, iters=1000000, ctx=None,
                    shuffle=_lib.SHUFFLE_FEISTEL, strict=None, placement=None):
    """Device-resident stage-2 loop on centred vertices (torch CUDA tensors).
    strict: True/False sets the context's strict guard (None keeps it).
    placement: "midpoint" (the reference's) or "optimal" (QEM optimal vertex,
    not in the reference) sets the context's placement (None keeps it).

    Returns (verts, faces, stats) as torch CUDA tensors and a dict."""
    import ctypes
    if ctx is None:
        ctx = _lib.Context(verts.device.index if verts.device.index is not None else torch.cuda.current_device(),
                           shuffle)
    verts = verts.contiguous()
    faces = faces.contiguous()
    if verts.dtype != torch.float32 or faces.dtype != torch.int32:
        raise RuntimeError("verts must be float32 and faces int32")
    if strict is not None:
        ctx.set_strict(strict)
    if placement is not None:
        ctx.set_placement(placement)
    L = ctx._L
    stream = torch.cuda.current_stream().cuda_stream
    oV, oF = ctypes.c_int(), ctypes.c_int()
    st = np.zeros(8, np.int64)
    rc = L.pamo_run_stage2(ctx.handle, stream, verts.data_ptr(), faces.data_ptr(), int(verts.shape[0]),
                           int(faces.shape[0]), float(scale), int(target), float(threshold), int(tolerance),
                           int(iters), ctypes.byref(oV), ctypes.byref(oF), st.ctypes.data)
    ctx.check(rc, "pamo_run_stage2")
    ov = torch.empty((oV.value, 3), dtype=torch.float32, device=verts.device)
    of = torch.empty((oF.value, 3), dtype=torch.int32, device=verts.device)
    ctx.check(L.pamo_run_fetch(ctx.handle, stream, ov.data_ptr(), of.data_ptr()), "pamo_run_fetch")
    stats = dict(rounds=int(st[0]), exit_reason=int(st[1]), collapses=int(st[2]), max_undo_passes=int(st[3]),
                 checks=int(st[4]), query_faces=int(st[5]), candidate_pairs=int(st[6]), hit_pairs=int(st[7]))
    rf = np.zeros(max(stats["rounds"], 1), np.int32)
    L.pamo_run_rounds(ctx.handle, rf.ctypes.data, rf.size)
    stats["round_faces"] = rf[:stats["rounds"]]
    return ov, of, stats


class PaMO(torch.nn.Module):
    """simp_cuda/pamo/__init__.py:18-178: stage 1 (remeshing, optional),
    stage 2 and stage 3 (safe projection toward the input mesh, optional).

    ``engine="device"`` (default) keeps the whole round loop on the GPU;
    ``engine="rounds"`` runs the reference's Python loop over
    ``CUDSP_Free.forward`` with torch compaction.  Both give identical
    results."""

    def __init__(self, input_mesh, use_stage1=True, use_stage3=True, verbose=True, engine="device",
                 shuffle=_lib.SHUFFLE_FEISTEL, strict=False, placement="midpoint"):
        super().__init__()
        self.use_stage1 = use_stage1
        self.use_stage3 = use_stage3
        self.verbose = verbose
        self.engine = engine
        if verbose:
            print("Stage1 : ", self.use_stage1)
            print("Stage3 : ", self.use_stage3)
        v, f = _as_vf(input_mesh)
        self.bbox = np.stack([v.min(0), v.max(0)])
        # the target of stage 3 (the reference keeps a copy of the input mesh)
        self.gt_vertices = np.array(v, np.float64)
        self.gt_faces = np.array(f, np.int32)
        self.pamo = _C.CUDSP_Free(shuffle)
        if strict and engine != "device":
            raise ValueError("strict=True needs engine='device' (the round kernels follow the reference)")
        self.pamo._ctx.set_strict(strict)
        if placement != "midpoint" and engine != "device":
            raise ValueError("placement=%r needs engine='device' (the round kernels follow the reference)" % placement)
        self.pamo._ctx.set_placement(placement)
        # stage 1's lattice (simp_cuda/pamo/__init__.py:49-52): the band and
        # margin stay those of R = 256 when run() lowers R for small targets
        self.R = 256
        self.band = 3 / self.R
        self.margin = self.band * 2 + 1
        self.target_faces = None
        self.last_stats = None
        self.last_stage2_seconds = None
        self.last_stage1_seconds = None
        self.last_stage3_seconds = None
        self.last_stage2_output = None
        if use_stage3:
            # one stage-3 system, reused by every run (simp_cuda/pamo/__init__.py:31-33)
            from . import safe_project
            self.config = safe_project.Stage3Config()
            self.system = safe_project.Stage3System(self.config, ctx=self.pamo._ctx)

    def run(self, points, triangles, ratio, tolerance=4, threshold=1e-3, iter=1000000, min_verts=10000000000):
        self.target_faces = max(int(ratio * len(triangles)), min_verts)
        if self.verbose:
            print("Target faces : {}".format(self.target_faces))
        # the reference's centring and scale (its torch/numpy float32
        # expressions, evaluated in one host pass with the same results)
        hv, _, scale, tris_mean = prepare_host(points.cpu().numpy(), triangles.cpu().numpy())
        verts = torch.from_numpy(hv).to(points.device)
        faces = triangles
        if self.use_stage1:
            from . import remesh
            self.R = remesh.resolution(self.target_faces)
            start1 = time.time()
            tris, tris_min, tris_max, tris_mean = remesh.normalize(points.cpu().numpy(), triangles.cpu().numpy(),
                                                                   self.band, self.margin)
            verts, faces = remesh.remesh(torch.from_numpy(tris).to(points.device), tris_min, tris_max, self.R,
                                         self.band, self.margin, ctx=self.pamo._ctx)
            torch.cuda.synchronize()
            self.last_stage1_seconds = time.time() - start1
            if self.verbose:
                print(f"Time for Remeshing: {self.last_stage1_seconds} sec")
            # stage 2's scale over the remeshed vertices (__init__.py:126)
            scale = _scale(verts)
        start = time.time()
        if self.engine == "device":
            verts, faces, stats = simplify_stage2(verts, faces, scale, self.target_faces, threshold, tolerance,
                                                  iter, ctx=self.pamo._ctx)
            if stats["exit_reason"] == 3 and self.verbose:
                print("Not enough edges available to be collapsed.")
        else:
            verts, faces, stats = self._run_rounds(verts, faces, scale, tolerance, threshold, iter)
        torch.cuda.synchronize()
        end = time.time()
        self.last_stats = stats
        self.last_stage2_seconds = end - start
        verts = verts.cpu().numpy() + tris_mean
        faces = faces.cpu().numpy()
        if self.verbose:
            print(f"Time for Simplification: {end - start} sec")
        self.last_stage2_output = (verts, faces)
        if self.use_stage3:
            # simp_cuda/pamo/__init__.py:163-176: 5 steps toward the input mesh
            from . import safe_project
            start3 = time.time()
            verts, faces = safe_project.process(self.gt_vertices, self.gt_faces, verts, faces, 5, system=self.system,
                                                config=self.config)
            self.last_stage3_seconds = time.time() - start3
        return verts, faces

    def _run_rounds(self, verts, faces, scale, tolerance, threshold, iters):
        """The reference's loop (simp_cuda/pamo/__init__.py:122-158) verbatim
        in behaviour, one CUDSP_Free.forward per round."""
        verts_undo = torch.empty(0, dtype=torch.int32, device=verts.device)
        n_verts_undo = 0
        count = 0
        is_stuck = 0
        init = True
        rounds = 0
        exit_reason = 4
        for it in range(iters):
            num_faces_prev = faces.shape[0]
            verts, faces, verts_occ, verts_map, verts_undo = self.pamo.forward(
                verts, faces, verts_undo, n_verts_undo, scale, threshold, is_stuck, init)
            init = False
            rounds += 1
            n_verts_undo = verts_undo.shape[0]
            verts = verts[verts_occ.view(-1).bool()]
            faces = faces[faces[:, 0] >= 0]
            faces[:, 0] = verts_map[faces[:, 0].long()].view(-1)
            faces[:, 1] = verts_map[faces[:, 1].long()].view(-1)
            faces[:, 2] = verts_map[faces[:, 2].long()].view(-1)
            num_faces_current = faces.shape[0]
            if num_faces_current <= self.target_faces or num_faces_current <= 10:
                exit_reason = 1 if num_faces_current <= self.target_faces else 2
                break
            if num_faces_current == num_faces_prev:
                count += 1
            else:
                count = 0
                is_stuck = 0
            if count >= 2:
                is_stuck = 1
            if count == tolerance:
                if self.verbose:
                    print("Not enough edges available to be collapsed.")
                exit_reason = 3
                break
        return verts, faces, dict(rounds=rounds, exit_reason=exit_reason)


class PaSP(torch.nn.Module):
    """Unguarded simplifier (simp_cuda/pamo/__init__.py:180-227): rounds of
    _C.CUDSP.forward with the same compaction as PaMO.run, until a round
    removes no face or `iter` rounds have run.  No centring; scale is the
    largest axis extent of the input."""

    def __init__(self):
        super().__init__()
        self.sp = _C.CUDSP()

    def run(self, points, triangles, threshold=0.001, iter=1000):
        verts = points
        faces = triangles
        scale = max(max(verts[:, 0].max() - verts[:, 0].min(), verts[:, 1].max() - verts[:, 1].min()),
                    verts[:, 2].max() - verts[:, 2].min())
        init = True
        self.rounds = 0
        for it in range(iter):
            num_faces = faces.shape[0]
            verts, faces, verts_occ, verts_map = self.sp.forward(verts, faces, float(scale), threshold, init)
            verts = verts[verts_occ.view(-1).bool()]
            faces = faces[faces[:, 0] >= 0]
            faces[:, 0] = verts_map[faces[:, 0].long()].view(-1)
            faces[:, 1] = verts_map[faces[:, 1].long()].view(-1)
            faces[:, 2] = verts_map[faces[:, 2].long()].view(-1)
            init = False
            self.rounds = it + 1
            if faces.shape[0] == num_faces:
                break
        return verts, faces


def simplify_host(points, faces, threshold=1e-3):
    """The Blender add-on's call (pamo_blender/__init__.py:55-60 over
    pamo_simplify): host arrays in and out, the guarded loop on the raw
    vertices until the face count stalls 4 times.  Returns (verts, faces)."""
    import numpy as np
    v = np.ascontiguousarray(points, np.float32)
    t = np.ascontiguousarray(faces, np.int32)
    vo = np.empty_like(v)
    to = np.empty_like(t)
    sh = np.array([v.shape[0], t.shape[0]], np.int32)
    _lib.lib().pamo_simplify(v.ctypes.data, t.ctypes.data, vo.ctypes.data, to.ctypes.data, sh.ctypes.data,
                             float(threshold))
    if sh[0] < 0:
        raise _lib.PamoError("pamo_simplify failed (see stderr)")
    return vo[:sh[0]].copy(), to[:sh[1]].copy()
