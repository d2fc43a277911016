"""pamo_run_stage2_noalloc, the variant of pamo_run_stage2 that allocates
nothing (working buffers carved from the library's fixed device region),
called through the C ABI: two calls in a row, the second on a smaller mesh,
each equal bit for bit to pamo_run_stage2 on the same input, and a mesh
above the size limit rejected before any GPU work."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from pamo_amd import _lib, synth

pytestmark = pytest.mark.gpu

MAX_FACES = 65536


def _two_spheres(level=3, seed=3):
    v1, f1 = synth.displaced_icosphere(level, seed)
    v2, f2 = synth.displaced_icosphere(level, seed + 1)
    v2 = v2 + np.array([0.9, 0.05, 0.0], np.float32)
    return np.concatenate([v1, v2]).astype(np.float32), np.concatenate([f1, f2 + v1.shape[0]]).astype(np.int32)


def _run(ctx, fn, verts, faces, scale, target):
    L = ctx._L
    dv = torch.from_numpy(verts).cuda()
    df = torch.from_numpy(faces).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    oV, oF = ctypes.c_int(), ctypes.c_int()
    st = np.zeros(8, np.int64)
    if fn == "noalloc":
        # outputs sized as the inputs, filled with a marker so a short write shows
        ov = torch.full((verts.shape[0], 3), 7.0, dtype=torch.float32, device="cuda")
        of = torch.full((faces.shape[0], 3), -7, dtype=torch.int32, device="cuda")
        rc = L.pamo_run_stage2_noalloc(ctx.handle, stream, dv.data_ptr(), df.data_ptr(), verts.shape[0],
                                       faces.shape[0], float(scale), int(target), 1e-3, 4, 1000000, ov.data_ptr(),
                                       of.data_ptr(), ctypes.byref(oV), ctypes.byref(oF), st.ctypes.data)
        ctx.check(rc, "pamo_run_stage2_noalloc")
        assert (ov[oV.value:] == 7.0).all() and (of[oF.value:] == -7).all()
        return ov[:oV.value].cpu().numpy(), of[:oF.value].cpu().numpy(), st
    rc = L.pamo_run_stage2(ctx.handle, stream, dv.data_ptr(), df.data_ptr(), verts.shape[0], faces.shape[0],
                           float(scale), int(target), 1e-3, 4, 1000000, ctypes.byref(oV), ctypes.byref(oF),
                           st.ctypes.data)
    ctx.check(rc, "pamo_run_stage2")
    ov = torch.empty((oV.value, 3), dtype=torch.float32, device="cuda")
    of = torch.empty((oF.value, 3), dtype=torch.int32, device="cuda")
    ctx.check(L.pamo_run_fetch(ctx.handle, stream, ov.data_ptr(), of.data_ptr()), "pamo_run_fetch")
    return ov.cpu().numpy(), of.cpu().numpy(), st


def test_noalloc_equals_run_stage2(cuda):
    big = synth.sdf_union(20, 115, 31)                 # 65,104 faces, near the limit
    assert 60000 < big[1].shape[0] <= MAX_FACES
    cases = [big, _two_spheres(), synth.noisy_torus(48, 20, 6)]   # then smaller ones (the second intersects)
    ref_ctx = _lib.Context(0)
    na_ctx = _lib.Context(0)
    for v, f in cases:
        verts, scale, _ = oracle.preprocess(v, f)
        target = int(0.01 * f.shape[0])
        print("noalloc case", f.shape, flush=True)
        rv, rf, rst = _run(ref_ctx, "stage2", verts, f, scale, target)
        nv, nf, nst = _run(na_ctx, "noalloc", verts, f, scale, target)
        print("  rounds", rst[0], nst[0], "faces", rf.shape[0], nf.shape[0], flush=True)
        assert np.array_equal(rst, nst)
        assert rv.tobytes() == nv.tobytes() and rf.tobytes() == nf.tobytes()


def test_noalloc_rejects_oversize_and_strict(cuda):
    ctx = _lib.Context(0)
    L = ctx._L
    v, f = synth.displaced_icosphere(2, 1)
    dv = torch.from_numpy(v).cuda()
    df = torch.from_numpy(f).cuda()
    out_v = torch.empty_like(dv)
    out_f = torch.empty_like(df)
    oV, oF = ctypes.c_int(), ctypes.c_int()
    # sizes above the limit are refused on the host (nothing is launched)
    rc = L.pamo_run_stage2_noalloc(ctx.handle, None, dv.data_ptr(), df.data_ptr(), v.shape[0], MAX_FACES + 1, 1.0,
                                   0, 1e-3, 4, 10, out_v.data_ptr(), out_f.data_ptr(), ctypes.byref(oV),
                                   ctypes.byref(oF), None)
    assert rc == -4 and "at most" in ctx.error()
    ctx.set_strict(True)
    rc = L.pamo_run_stage2_noalloc(ctx.handle, None, dv.data_ptr(), df.data_ptr(), v.shape[0], f.shape[0], 1.0,
                                   0, 1e-3, 4, 10, out_v.data_ptr(), out_f.data_ptr(), ctypes.byref(oV),
                                   ctypes.byref(oF), None)
    assert rc == -4 and "strict" in ctx.error()
