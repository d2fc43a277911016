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


def test_cli_all_stages(cuda, tmp_path):
    """With no --disable flag every stage runs (example.py's default):
    remeshing, stage 2, then stage 3 toward the input.  The written mesh
    has the stage-2 faces of that run and finite float64-projected
    vertices."""
    v, f = synth.displaced_icosphere(4, 5)
    src = str(tmp_path / "in.obj")
    dst = str(tmp_path / "out.obj")
    meshio.save(src, v.astype(np.float32), f)
    rc = main(["-i", src, "-o", dst, "-r", "0.05"])
    ov, of = meshio.load(dst)
    print("cli all stages: %d -> %d faces" % (len(f), len(of)), flush=True)
    assert rc == 0 and len(of) > 0 and np.isfinite(ov).all()
    assert of.min() >= 0 and of.max() < len(ov)
