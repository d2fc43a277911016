"""C-ABI and host-side checks that need no GPU."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pamo_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pamo_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from pamo_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n


def test_library_is_gfx950_code_object():
    """The shared library carries a gfx950 device code object."""
    from pamo_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(monkeypatch):
    from pamo_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libpamo_hip.so")
    with pytest.raises(_lib.PamoError):
        _lib.lib()


def test_input_checks_match_reference_messages():
    import torch
    from pamo_amd._C import _check_input
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        _check_input(torch.zeros(3, 3), "points")


def test_pamo_without_gpu_fails_loudly():
    """Every stage runs on the device: without a GPU, PaMO refuses to build
    instead of falling back to the host."""
    import numpy as np
    import torch
    from pamo_amd import PaMO
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    for kw in ({}, {"use_stage1": False}, {"use_stage1": False, "use_stage3": False}):
        with pytest.raises(RuntimeError):
            PaMO((v, f), verbose=False, **kw)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_preprocess_matches_reference_expression():
    """tris_mean / centring / scale of simp_cuda/pamo/__init__.py:60-64,117,126
    in numpy equal the torch expressions the reference evaluates."""
    import numpy as np
    import torch
    from pamo_amd import preprocess_numpy, _preprocess, _scale
    from pamo_amd import synth
    v, f = synth.displaced_icosphere(3, 2)
    verts, scale, tm = preprocess_numpy(v, f)
    tm2 = _preprocess(torch.from_numpy(v), torch.from_numpy(f))
    assert tm.tobytes() == tm2.tobytes()
    vt = torch.from_numpy(v) - torch.from_numpy(tm2)
    assert vt.numpy().tobytes() == verts.tobytes()
    assert _scale(vt) == scale


@pytest.mark.parametrize("seed", range(6))
def test_prepare_host_matches_numpy(seed):
    """pamo_prepare_host (one pass in the library) equals the numpy
    expressions of the centring and scale bit for bit: random meshes of
    every size class with offsets and scales far from 1, and ragged index
    use (unreferenced vertices, repeated corners)."""
    import numpy as np
    from pamo_amd import batch, preprocess_numpy
    rng = np.random.default_rng(seed)
    for _ in range(6):
        V = int(rng.integers(1, 30000))
        F = int(rng.integers(1, 60000))
        p = (rng.standard_normal((V, 3)) * rng.uniform(1e-4, 1e4) + rng.uniform(-1e3, 1e3, 3)).astype(np.float32)
        f = rng.integers(0, V, (F, 3)).astype(np.int32)
        v1, t1, s1, m1 = batch.prepare_host(p, f)
        v2, s2, m2 = preprocess_numpy(p, f)
        assert v1.tobytes() == v2.tobytes()
        assert m1.tobytes() == m2.astype(np.float32).tobytes()
        assert s1 == s2
        assert t1.dtype == np.int32 and np.array_equal(t1, f)


def test_prepare_host_signed_zero_and_single_face():
    """Both sums start from +0 as numpy's add.reduce does (-0.0 terms add
    up to +0.0)."""
    import numpy as np
    from pamo_amd import batch, preprocess_numpy
    p = np.array([[-0.0, 1.0, 2.0], [-0.0, 3.0, 4.0], [-0.0, 5.0, 7.0]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    v1, _, s1, m1 = batch.prepare_host(p, f)
    v2, s2, m2 = preprocess_numpy(p, f)
    assert m1.tobytes() == m2.astype(np.float32).tobytes()
    assert v1.tobytes() == v2.tobytes() and s1 == s2
