"""The example.py-style command line (python -m pamo_amd): file in, PaMO
stage 2 on the GPU, file out; the written mesh equals the CPU oracle's run
of the same input, bit for bit after the float32 round trip of the file."""
import numpy as np
import pytest

import oracle
from pamo_amd import meshio, synth
from pamo_amd.__main__ import main

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ext", [".obj", ".ply"])
def test_cli_output_matches_oracle(cuda, tmp_path, ext):
    v, f = synth.sdf_union(16, 64, 21)
    v = v.astype(np.float32)
    src = str(tmp_path / ("in" + ext))
    dst = str(tmp_path / ("out" + ext))
    meshio.save(src, v, f)
    ratio = 0.02
    rc = main(["-i", src, "-o", dst, "-r", str(ratio), "--disable_stage1", "--disable_stage3", "--verify"])
    ov, of = meshio.load(dst)
    verts, scale, mean = oracle.preprocess(v, f)
    wv, wf, st = oracle.run(verts, f, scale, int(ratio * f.shape[0]))
    print("cli", ext, f.shape, st, flush=True)
    assert np.array_equal(of, wf)
    assert np.array_equal(ov.astype(np.float32), (wv + mean).astype(np.float32))
    assert rc == 0
