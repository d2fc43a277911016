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


def test_stage_flags_raise():
    import numpy as np
    from pamo_amd import PaMO
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    with pytest.raises(NotImplementedError):
        PaMO((v, f))
    with pytest.raises(NotImplementedError):
        PaMO((v, f), use_stage1=False)


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
