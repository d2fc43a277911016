"""Drop-in for the reference extension module ``pamo._C``.

``CUDSP_Free`` mirrors simp_cuda/src/pybind.cpp:16-103: same constructor,
same ``forward`` arguments, same input checks and error messages, the same
five freshly allocated output tensors.  The work runs in libpamo_hip.so on
the current torch stream.
"""
import torch

from . import _lib


def _check_input(t, name):
    # pybind.cpp:8-14 (AT_ASSERTM raises RuntimeError)
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError("%s must be a CUDA tensor" % name)
    if not t.is_contiguous():
        raise RuntimeError("%s must be contiguous" % name)


def _stream_handle():
    return torch.cuda.current_stream().cuda_stream


class CUDSP_Free:
    """Guarded stage-2 simplifier context (``pamo._C.CUDSP_Free``)."""

    def __init__(self, shuffle=_lib.SHUFFLE_FEISTEL):
        self._ctx = _lib.Context(torch.cuda.current_device(), shuffle)

    def forward(self, points, triangles, verts_undo, n_verts_undo, scale, threshold, is_stuck, init):
        _check_input(points, "points")
        _check_input(triangles, "triangles")
        _check_input(verts_undo, "verts_undo")
        if points.dtype != torch.float32:
            raise RuntimeError("points type must match the pamo class")
        if triangles.dtype != torch.int32:
            raise RuntimeError("triangles type must match the pamo class")
        n_verts_undo = int(n_verts_undo)
        if n_verts_undo > verts_undo.numel():
            raise RuntimeError("n_verts_undo exceeds verts_undo")
        if n_verts_undo and verts_undo.dtype != torch.int32:
            raise RuntimeError("verts_undo must be int32")
        n_pts = int(points.shape[0])
        n_tris = int(triangles.shape[0])
        L = self._ctx._L
        stream = _stream_handle()
        rc = L.pamo_free_forward(self._ctx.handle, stream, points.data_ptr(), triangles.data_ptr(),
                                 verts_undo.data_ptr() if n_verts_undo else None, n_verts_undo, n_pts, n_tris,
                                 float(scale), float(threshold), int(bool(is_stuck)), int(bool(init)))
        self._ctx.check(rc, "CUDSP_Free.forward")
        import ctypes
        nv, nf, nu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._ctx.check(L.pamo_free_outputs(self._ctx.handle, None, None, None, None, None, ctypes.byref(nv),
                                            ctypes.byref(nf), ctypes.byref(nu)), "outputs")
        dev = points.device
        verts = torch.empty((nv.value, 3), dtype=torch.float32, device=dev)
        tris = torch.empty((nf.value, 3), dtype=torch.int32, device=dev)
        occ = torch.empty((nv.value, 1), dtype=torch.int32, device=dev)
        vmap = torch.empty((nv.value, 1), dtype=torch.int32, device=dev)
        undo = torch.empty((nu.value,), dtype=torch.int32, device=dev)
        self._ctx.check(L.pamo_free_copy_outputs(self._ctx.handle, stream, verts.data_ptr(), tris.data_ptr(),
                                                 occ.data_ptr(), vmap.data_ptr(), undo.data_ptr()), "copy outputs")
        return verts, tris, occ, vmap, undo

    # debug exports for per-kernel parity (not part of the reference API)
    _DEBUG = dict(fan_off=(5, "int32"), fan=(6, "int32"), edges=(7, "int32"), Q=(8, "float32"),
                  edge_cost=(9, "uint32"), tri_min=(10, "uint64"), collapsed=(11, "int32"), S=(12, "uint8"),
                  orig_tris=(13, "int32"), table=(14, "uint8"))

    def debug(self, name):
        import numpy as np
        which, dt = self._DEBUG[name]
        L = self._ctx._L
        stream = _stream_handle()
        n = self._ctx.check(L.pamo_debug_copy(self._ctx.handle, stream, which, None, 0), "debug size")
        a = np.empty(n // np.dtype(dt).itemsize, dt)
        if n:
            self._ctx.check(L.pamo_debug_copy(self._ctx.handle, stream, which, a.ctypes.data, n), "debug copy")
        if name in ("orig_tris",):
            a = a.reshape(-1, 3)
        elif name == "edges":
            a = a.reshape(-1, 2)
        elif name == "Q":
            a = a.reshape(-1, 16)
        return a


class CUDSP:
    """Unguarded simplifier context (``pamo._C.CUDSP``, simp_cuda/src/pybind.cpp:
    119-183 over CUSimp::forward, simp_cuda/src/cusimp.cu:587-720): one
    collapse round without the self-intersection guard, the skinny-face
    restore, invalid-face removal or line-edge collapse."""

    def __init__(self, shuffle=_lib.SHUFFLE_FEISTEL):
        self._ctx = _lib.Context(torch.cuda.current_device(), shuffle)

    def forward(self, points, triangles, scale, threshold, init):
        _check_input(points, "points")
        _check_input(triangles, "triangles")
        if points.dtype != torch.float32:
            raise RuntimeError("points type must match the sp class")
        if triangles.dtype != torch.int32:
            raise RuntimeError("triangles type must match the sp class")
        n_pts = int(points.shape[0])
        n_tris = int(triangles.shape[0])
        L = self._ctx._L
        stream = _stream_handle()
        rc = L.pamo_sp_forward(self._ctx.handle, stream, points.data_ptr(), triangles.data_ptr(), n_pts, n_tris,
                               float(scale), float(threshold), int(bool(init)))
        self._ctx.check(rc, "CUDSP.forward")
        import ctypes
        nv, nf, nu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._ctx.check(L.pamo_free_outputs(self._ctx.handle, None, None, None, None, None, ctypes.byref(nv),
                                            ctypes.byref(nf), ctypes.byref(nu)), "outputs")
        dev = points.device
        verts = torch.empty((nv.value, 3), dtype=torch.float32, device=dev)
        tris = torch.empty((nf.value, 3), dtype=torch.int32, device=dev)
        occ = torch.empty((nv.value, 1), dtype=torch.int32, device=dev)
        vmap = torch.empty((nv.value, 1), dtype=torch.int32, device=dev)
        self._ctx.check(L.pamo_free_copy_outputs(self._ctx.handle, stream, verts.data_ptr(), tris.data_ptr(),
                                                 occ.data_ptr(), vmap.data_ptr(), None), "copy outputs")
        return verts, tris, occ, vmap

    # the same debug exports as CUDSP_Free (fans, edges, quadrics, costs, claims)
    _DEBUG = CUDSP_Free._DEBUG
    debug = CUDSP_Free.debug
