"""Parity hooks over libpamo_hip.so that the reference has no counterpart for:
the full candidate-pair set of one self-intersection check."""
import numpy as np
import torch

from . import _lib


def candidate_pairs(points, tris, ctx=None):
    """All ordered (query, current) pairs of alive faces whose float AABBs
    overlap, with the reference narrow-phase verdict, computed by the GPU
    LBVH (simp_cuda/src/bvh/self_intersect.cuh:151-392 over every face).

    points/tris: torch CUDA tensors.  Returns an int32 array of rows
    (query, current, hit) sorted by (query, current)."""
    if ctx is None:
        ctx = _lib.Context(torch.cuda.current_device())
    points = points.contiguous()
    tris = tris.contiguous()
    L = ctx._L
    stream = torch.cuda.current_stream().cuda_stream
    n = ctx.check(L.pamo_debug_pairs(ctx.handle, stream, points.data_ptr(), tris.data_ptr(), points.shape[0],
                                     tris.shape[0], None, 0), "debug pairs")
    out = np.empty((max(n, 0), 3), np.int32)
    if n:
        ctx.check(L.pamo_debug_pairs(ctx.handle, stream, points.data_ptr(), tris.data_ptr(), points.shape[0],
                                     tris.shape[0], out.ctypes.data, n), "debug pairs")
    order = np.lexsort((out[:, 1], out[:, 0]))
    return out[order]
