"""Stage 1 (remeshing) on the GPU: the two calls PaMO.remesh makes
(simp_cuda/pamo/__init__.py:75-93).

``get_sdf(tris, R, band)`` stands for ``torchcumesh2sdf.get_sdf`` and
``VolumeToMesh`` for ``pdmc.DMC``; the reference calls both packages but does
not ship them, so the library defines them (pamo_mesh_to_sdf and
pamo_volume_to_mesh in include/pamo_hip.h):

* the volume is an R^3 lattice, sample (i, j, k) at ((i + 0.5) / R, ...),
  holding the exact distance to the nearest triangle within ``band`` and
  +-band beyond it, positive outside (samples connected to the lattice
  boundary through samples at least 0.75 / R from the surface; samples
  nearer than that take the side of their nearest triangle, whose
  orientation the outside and inside samples vote on);
* the extraction returns the volume's zero level set as a closed,
  consistently (outward) oriented 2-manifold, vertices in the lattice's
  [0, 1]^3 coordinates.

``remesh`` runs the reference's sequence: normalisation of the triangles
into [0, 1]^3 with a band margin, the volume, the shift by 0.9 / R (a shell
0.9 lattice steps outside the input), the extraction, and the map back.
"""
import ctypes

import numpy as np
import torch

from . import _lib

R_DEFAULT = 256
MAX_R = 512


def _bind(L):
    if getattr(L, "_pamo_remesh_bound", False):
        return L
    P, I = ctypes.c_void_p, ctypes.c_int
    L.pamo_mesh_to_sdf.argtypes = [P, P, P, I, I, ctypes.c_float, P]
    L.pamo_mesh_to_sdf.restype = I
    L.pamo_volume_to_mesh.argtypes = [P, P, P, I, ctypes.POINTER(I), ctypes.POINTER(I)]
    L.pamo_volume_to_mesh.restype = I
    L.pamo_volume_mesh_fetch.argtypes = [P, P, P, P]
    L.pamo_volume_mesh_fetch.restype = I
    L._pamo_remesh_bound = True
    return L


def _ctx_for(t, ctx):
    if ctx is not None:
        return ctx
    dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
    return _lib.Context(dev)


def get_sdf(tris, R, band, ctx=None):
    """Signed distance lattice of triangles ([n, 3, 3] float32 CUDA tensor in
    [0, 1]^3): an [R, R, R] float32 CUDA tensor indexed [k, j, i]."""
    if not (isinstance(tris, torch.Tensor) and tris.is_cuda):
        raise RuntimeError("tris must be a CUDA tensor")
    tris = tris.to(torch.float32).contiguous()
    ctx = _ctx_for(tris, ctx)
    L = _bind(ctx._L)
    vol = torch.empty((R, R, R), dtype=torch.float32, device=tris.device)
    n = tris.numel() // 9
    rc = L.pamo_mesh_to_sdf(ctx.handle, torch.cuda.current_stream().cuda_stream, tris.data_ptr(), int(n), int(R),
                            float(band), vol.data_ptr())
    ctx.check(rc, "pamo_mesh_to_sdf")
    return vol


class VolumeToMesh:
    """Zero level set of an [R, R, R] volume (positive: outside) as
    (verts [V, 3] float32, faces [F, 3] int32) CUDA tensors, vertices in
    [0, 1]^3 lattice coordinates.  Called like the reference's
    ``DMC(d, return_quads=False)``."""

    def __init__(self, ctx=None):
        self.ctx = ctx

    def __call__(self, d, return_quads=False):
        if return_quads:
            raise NotImplementedError("the extraction returns triangles only")
        if not (isinstance(d, torch.Tensor) and d.is_cuda) or d.dim() != 3 or len(set(d.shape)) != 1:
            raise RuntimeError("d must be a cubic CUDA volume")
        d = d.to(torch.float32).contiguous()
        R = d.shape[0]
        ctx = _ctx_for(d, self.ctx)
        self.ctx = ctx
        L = _bind(ctx._L)
        stream = torch.cuda.current_stream().cuda_stream
        nv, nf = ctypes.c_int(), ctypes.c_int()
        ctx.check(L.pamo_volume_to_mesh(ctx.handle, stream, d.data_ptr(), int(R), ctypes.byref(nv), ctypes.byref(nf)),
                  "pamo_volume_to_mesh")
        v = torch.empty((nv.value, 3), dtype=torch.float32, device=d.device)
        f = torch.empty((nf.value, 3), dtype=torch.int32, device=d.device)
        ctx.check(L.pamo_volume_mesh_fetch(ctx.handle, stream, v.data_ptr(), f.data_ptr()), "pamo_volume_mesh_fetch")
        return v, f


def normalize(points, triangles, band, margin):
    """PaMO.preprocess_mesh (simp_cuda/pamo/__init__.py:57-68) on host
    arrays: (tris [F, 3, 3] float32 in [0, 1]^3, tris_min, tris_max,
    tris_mean)."""
    tris = np.asarray(points, np.float32)[np.asarray(triangles)]
    tris_mean = tris.mean(axis=1).mean(axis=0)
    tris = tris - tris_mean
    tris_min = tris.min(0).min(0)
    tris = tris - tris_min
    tris_max = tris.max()
    tris = (tris / tris_max + band) / margin
    return tris.astype(np.float32), tris_min, tris_max, tris_mean


def remesh(tris, tris_min, tris_max, R, band, margin, ctx=None, vol2mesh=None):
    """PaMO.remesh (simp_cuda/pamo/__init__.py:75-93): the shell 0.9 / R
    outside the normalised triangles (a CUDA tensor), mapped back to the
    centred input frame.  Returns (verts float32, faces int32) CUDA tensors."""
    d = get_sdf(tris, R, band, ctx=ctx)
    d = d - 0.9 / R
    vol2mesh = vol2mesh or VolumeToMesh(ctx)
    v, f = vol2mesh(d)
    # lattice [0, 1]^3 -> normalised -> centred input frame
    v = (v * margin - band) * float(tris_max) + torch.from_numpy(np.asarray(tris_min, np.float32)).to(v.device)
    return v.contiguous(), f


def resolution(target_faces):
    """The reference's lattice resolution by face target (__init__.py:106-110)."""
    R = R_DEFAULT
    if target_faces <= 1000:
        R = 128
    if target_faces <= 50:
        R = 64
    return R
